// lua_host.cpp -- a dort Lua host with the GPU renderers (TEST INFRASTRUCTURE ONLY).
//
// Built by oracle/ref.mk (target `lua`) into oracle/_ref/dort_lua from dort's own
// sources, the Lua 5.3 VM and zlib vendored with it (extern/lua, extern/zlib), and
// dort_amd/host/{gpu_path_renderer,lua_gpu}.cpp over libdortgpu.so.  It sets up a Lua
// state the way dort's main does (main.cpp:62-140, 142-229): the standard libraries,
// zlib, the `dort` table of dort's modules, dort's Lua extensions, a CtxG with a thread
// pool; then it loads the Lua module "dort_gpu" and its Lua half (--prelude), and runs
// the script.
//
//   dort_lua [-t threads] [--lua-dir DIR] [--prelude FILE] [--] script.lua [args...]
//
// dort embeds its .lua files as bytecode generated at build time (lua_sources.sh) and
// finds them with a builtin searcher (lua.cpp:28-68).  This host does not install that
// searcher; with --lua-dir, `require` reads dort's .lua files from DIR through Lua's
// standard file searcher, and DIR/dort/builder_ext.lua and geometry_ext.lua run where
// dort runs their embedded copies (main.cpp:203-204).  Without --lua-dir only dort's
// C modules are there (enough for scripts that call dort.builder etc. directly).
//
// Exit status as dort's: 0 ok, 1 Lua error, 2 uncaught C++ exception, 3 bad usage.
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "dort/ctx.hpp"
#include "dort/lua.hpp"
#include "dort/lua_builder.hpp"
#include "dort/lua_camera.hpp"
#include "dort/lua_chrono.hpp"
#include "dort/lua_env.hpp"
#include "dort/lua_filter.hpp"
#include "dort/lua_geometry.hpp"
#include "dort/lua_grid.hpp"
#include "dort/lua_image.hpp"
#include "dort/lua_light.hpp"
#include "dort/lua_material.hpp"
#include "dort/lua_math.hpp"
#include "dort/lua_nbt.hpp"
#include "dort/lua_render.hpp"
#include "dort/lua_sampler.hpp"
#include "dort/lua_shape.hpp"
#include "dort/lua_spectrum.hpp"
#include "dort/lua_stats.hpp"
#include "dort/lua_texture.hpp"
#include "dort/stats.hpp"
#include "dort/thread_pool.hpp"
#include "../dort_amd/host/lua_gpu.hpp"

namespace {

const char USAGE[] =
    "usage: dort_lua [-t threads] [--lua-dir DIR] [--prelude FILE] [--] script.lua [args...]\n";

int traceback(lua_State* l) {
  if (!lua_isstring(l, 1)) return 1;
  luaL_traceback(l, l, lua_tostring(l, 1), 1);
  return 1;
}

// Runs a Lua file under the traceback handler; false (message printed) on error.
bool run_file(lua_State* l, const char* path, const char* what) {
  lua_pushcfunction(l, traceback);
  int h = lua_gettop(l);
  int st = luaL_loadfile(l, path);
  if (st == LUA_OK) st = lua_pcall(l, 0, 0, h);
  if (st != LUA_OK) {
    std::fprintf(stderr, "%s error: %s\n", what, lua_tostring(l, -1));
    lua_settop(l, h - 1);
    return false;
  }
  lua_settop(l, h - 1);
  return true;
}

bool setup(lua_State* l, const std::string& lua_dir) {
  luaL_requiref(l, "_G", luaopen_base, 1);
  luaL_requiref(l, LUA_DBLIBNAME, luaopen_debug, 1);
  luaL_requiref(l, LUA_IOLIBNAME, luaopen_io, 1);
  luaL_requiref(l, LUA_LOADLIBNAME, luaopen_package, 1);
  luaL_requiref(l, LUA_MATHLIBNAME, luaopen_math, 1);
  luaL_requiref(l, LUA_OSLIBNAME, luaopen_os, 1);
  luaL_requiref(l, LUA_STRLIBNAME, luaopen_string, 1);
  luaL_requiref(l, LUA_TABLIBNAME, luaopen_table, 1);
  luaL_requiref(l, LUA_COLIBNAME, luaopen_coroutine, 1);
  luaL_requiref(l, "zlib", luaopen_zlib, 1);
  lua_settop(l, 0);

  // the `dort` table with each module also under package.loaded["dort.<name>"]
  static const struct { const char* name; lua_CFunction open; } mods[] = {
      {"builder", dort::lua_open_builder},   {"camera", dort::lua_open_camera},
      {"chrono", dort::lua_open_chrono},     {"env", dort::lua_open_env},
      {"filter", dort::lua_open_filter},     {"geometry", dort::lua_open_geometry},
      {"grid", dort::lua_open_grid},         {"image", dort::lua_open_image},
      {"light", dort::lua_open_light},       {"material", dort::lua_open_material},
      {"math", dort::lua_open_math},         {"nbt", dort::lua_open_nbt},
      {"render", dort::lua_open_render},     {"sampler", dort::lua_open_sampler},
      {"shape", dort::lua_open_shape},       {"spectrum", dort::lua_open_spectrum},
      {"stats", dort::lua_open_stats},       {"texture", dort::lua_open_texture}};
  lua_newtable(l);
  for (const auto& m : mods) {
    std::string full = std::string("dort.") + m.name;
    luaL_requiref(l, full.c_str(), m.open, 0);
    lua_setfield(l, -2, m.name);
  }
  lua_pushvalue(l, -1);
  lua_setglobal(l, "dort");
  // require "dort" gives the table, as dort's searcher does (lua.cpp:39-45)
  lua_getglobal(l, "package");
  lua_getfield(l, -1, "loaded");
  lua_pushvalue(l, -3);
  lua_setfield(l, -2, "dort");
  lua_pop(l, 1);
  lua_pushstring(l, lua_dir.empty() ? "" : (lua_dir + "/?.lua").c_str());
  lua_setfield(l, -2, "path");
  lua_pushstring(l, "");
  lua_setfield(l, -2, "cpath");
  lua_settop(l, 0);

  if (!lua_dir.empty()) {
    if (!run_file(l, (lua_dir + "/dort/builder_ext.lua").c_str(), "builder_ext")) return false;
    if (!run_file(l, (lua_dir + "/dort/geometry_ext.lua").c_str(), "geometry_ext")) return false;
  }

  luaL_requiref(l, "dort_gpu", dort_amd::luaopen_dort_gpu, 1);
  lua_settop(l, 0);
  return true;
}

int run(int argc, char** argv) {
  uint32_t threads = 1;
  std::string lua_dir, prelude;
  int i = 1;
  for (; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--") { ++i; break; }
    if ((a == "-t" || a == "--threads") && i + 1 < argc) {
      if (std::sscanf(argv[++i], "%u", &threads) != 1 || threads == 0) { std::fputs(USAGE, stderr); return 3; }
    } else if (a == "--lua-dir" && i + 1 < argc) {
      lua_dir = argv[++i];
    } else if (a == "--prelude" && i + 1 < argc) {
      prelude = argv[++i];
    } else if (a == "-h" || a == "--help") {
      std::fputs(USAGE, stdout);
      return 0;
    } else if (!a.empty() && a[0] == '-') {
      std::fputs(USAGE, stderr);
      return 3;
    } else {
      break;
    }
  }
  if (i >= argc) { std::fputs(USAGE, stderr); return 3; }
  const char* script = argv[i++];

  dort::stat_init_global();
  dort::stat_init_thread();
  dort::CtxG ctx;
  ctx.pool = std::make_shared<dort::ThreadPool>();
  for (int k = i; k < argc; ++k) ctx.argv.push_back(argv[k]);
  ctx.pool->start(threads);

  lua_State* l = luaL_newstate();
  dort::lua_set_ctx(l, &ctx);
  int status = 1;
  try {
    if (setup(l, lua_dir) && (prelude.empty() || run_file(l, prelude.c_str(), "prelude"))) {
      // the script's arguments as the global `arg` (arg[0] is the script)
      lua_createtable(l, argc - i, 1);
      lua_pushstring(l, script);
      lua_rawseti(l, -2, 0);
      for (int k = i; k < argc; ++k) {
        lua_pushstring(l, argv[k]);
        lua_rawseti(l, -2, k - i + 1);
      }
      lua_setglobal(l, "arg");
      status = run_file(l, script, "Runtime") ? 0 : 1;
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "Uncaught std::exception: %s\n", e.what());
    status = 2;
  } catch (...) {
    std::fprintf(stderr, "Uncaught unknown exception\n");
    status = 2;
  }
  lua_close(l);
  ctx.pool->stop();
  return status;
}

}  // namespace

int main(int argc, char** argv) { return run(argc, argv); }
