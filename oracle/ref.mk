# Builds oracle/_ref/libdortref.so: dort's own pt-path sources, its PLY reader
# (with the rply C parser it links), image I/O (stb) and convergence test, compiled
# where they lie under /root/reference, linked with ref_driver.cpp and
# ref_io_driver.cpp (TEST INFRASTRUCTURE ONLY; the output stays out of git).
#
#   make -f oracle/ref.mk            (needs /root/reference; nothing is copied)
#
# dort's Tupfile.lua builds with clang++ and its "fast" flags (-O3 -DNDEBUG
# -DDORT_DISABLE_STAT); the compiler matters: clang evaluates call arguments left
# to right, which fixes the order of the two draws in Sampler::random_2d.  The
# sources predate C++17 and rely on <functional>, <string> and <stdexcept>
# arriving through other headers; -include supplies those standard headers.
REF ?= /root/reference
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
REF_CXX ?= $(firstword $(wildcard /opt/rocm/llvm/bin/clang++) clang++)
CXXFLAGS := -O3 -std=c++14 -fPIC -DDORT_DISABLE_STAT -DNDEBUG \
  -include functional -include string -include stdexcept \
  -I$(REF)/include -I$(REF)/extern -I$(REF)/extern/stb -I$(REF)/extern/rply \
  -I$(REF)/extern/lua -I$(REF)/extern/zlib -w
REF_CC ?= $(firstword $(wildcard /opt/rocm/llvm/bin/clang) clang)

SRCS := path_renderer renderer film bvh bvh_primitive mesh_bvh_primitive triangle \
  primitive scene bsdf lambert_material microfacet rough_dielectric_material \
  dielectric_material fresnel diffuse_light shape light environment_light \
  infinite_light piecewise_distrib_2d discrete_distrib_1d monte_carlo \
  pinhole_camera camera transform random_sampler sampler sampled_filter filter \
  stats thread_pool box geometry math image image_texture texture_maps \
  list_primitive mat triangle_shape triangle_shape_primitive mesh_triangle_primitive \
  ply_mesh convergence_test light_renderer bdpt_renderer
OBJS := $(addprefix $(OUT)/obj/,$(addsuffix .o,$(SRCS))) $(OUT)/obj/rply.o $(OUT)/obj/ref_io_driver.o

all: $(OUT)/libdortref.so

$(OUT)/obj/%.o: $(REF)/src/dort/%.cpp
	@mkdir -p $(OUT)/obj
	$(REF_CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/ref_driver.o: $(HERE)ref_driver.cpp $(HERE)../include/dortgpu.h
	@mkdir -p $(OUT)/obj
	$(REF_CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/ref_io_driver.o: $(HERE)ref_io_driver.cpp
	@mkdir -p $(OUT)/obj
	$(REF_CXX) $(CXXFLAGS) -c $< -o $@

# rply (extern/rply, the PLY parser dort links) is C
$(OUT)/obj/rply.o: $(REF)/extern/rply/rply.c
	@mkdir -p $(OUT)/obj
	$(REF_CC) -O2 -fPIC -w -I$(REF)/extern/rply -c $< -o $@

$(OUT)/libdortref.so: $(OBJS) $(OUT)/obj/ref_driver.o
	$(REF_CXX) -shared -o $@ $^ -lpthread

# dort with dort_amd::GpuPathRenderer (dort_amd/host) in its Renderer seam, over
# libdortgpu.so; built when that library exists (target `gpu`).
GPU_LIB := $(HERE)../dort_amd/libdortgpu.so
HOST := $(HERE)../dort_amd/host

$(OUT)/obj/gpu_path_renderer.o: $(HOST)/gpu_path_renderer.cpp $(HOST)/gpu_path_renderer.hpp $(HERE)../include/dortgpu.h
	@mkdir -p $(OUT)/obj
	$(REF_CXX) $(CXXFLAGS) -I$(HERE)../include -c $< -o $@

$(OUT)/obj/ref_gpu_glue.o: $(HERE)ref_gpu_glue.cpp $(HOST)/gpu_path_renderer.hpp $(HERE)../include/dortgpu.h
	@mkdir -p $(OUT)/obj
	$(REF_CXX) $(CXXFLAGS) -I$(HERE)../include -c $< -o $@

$(OUT)/libdortref_gpu.so: $(OBJS) $(OUT)/obj/ref_driver.o $(OUT)/obj/gpu_path_renderer.o $(OUT)/obj/ref_gpu_glue.o $(GPU_LIB)
	$(REF_CXX) -shared -o $@ $(filter %.o,$^) -L$(dir $(GPU_LIB)) -ldortgpu -Wl,-rpath,'$$ORIGIN/../../dort_amd' -lpthread

gpu: $(OUT)/libdortref_gpu.so

# dort's Lua front-end with the GPU renderers (target `lua`): oracle/_ref/dort_lua, a
# host program (lua_host.cpp) over every dort source but main.cpp, the Lua 5.3 VM and
# zlib vendored with dort (extern/lua, extern/zlib, compiled as dort's Tupfile.lua:52-61
# lists them) and dort_amd/host/lua_gpu.cpp, over libdortgpu.so.  dort's lua.cpp also
# holds the builtin searcher over the Lua bytecode dort's build generates
# (lua_sources.sh); it is compiled into a section of its own, nothing here calls it,
# and --gc-sections drops it with its reference to the generated table, so no
# generated file is needed (the host reads dort's .lua files as text instead).
DORT_ALL := $(filter-out main,$(basename $(notdir $(wildcard $(REF)/src/dort/*.cpp))))
LUA_VM := $(basename $(notdir $(wildcard $(REF)/extern/lua/*.c)))
ZLIB := $(basename $(notdir $(wildcard $(REF)/extern/zlib/*.c)))
LUA_OBJS := $(addprefix $(OUT)/obj/,$(addsuffix .o,$(DORT_ALL))) \
  $(addprefix $(OUT)/obj/luavm/,$(addsuffix .o,$(LUA_VM))) \
  $(addprefix $(OUT)/obj/zlib/,$(addsuffix .o,$(ZLIB))) \
  $(OUT)/obj/rply.o $(OUT)/obj/gpu_path_renderer.o $(OUT)/obj/lua_gpu.o $(OUT)/obj/lua_host.o
# (zlib's own configuration macro Z_HAVE_UNISTD_H declares read/write/close for its gz*
# files, which today's clang no longer accepts undeclared)
CFLAGS_LUA := -O2 -std=c11 -fPIC -w -I$(REF)/extern/lua -I$(REF)/extern/zlib

$(OUT)/obj/lua.o: CXXFLAGS += -ffunction-sections -fdata-sections

$(OUT)/obj/luavm/%.o: $(REF)/extern/lua/%.c
	@mkdir -p $(OUT)/obj/luavm
	$(REF_CC) $(CFLAGS_LUA) -c $< -o $@

$(OUT)/obj/zlib/%.o: $(REF)/extern/zlib/%.c
	@mkdir -p $(OUT)/obj/zlib
	$(REF_CC) $(CFLAGS_LUA) -DZ_HAVE_UNISTD_H -c $< -o $@

$(OUT)/obj/lua_gpu.o: $(HOST)/lua_gpu.cpp $(HOST)/lua_gpu.hpp $(HOST)/gpu_path_renderer.hpp $(HERE)../include/dortgpu.h
	@mkdir -p $(OUT)/obj
	$(REF_CXX) $(CXXFLAGS) -I$(HERE)../include -c $< -o $@

$(OUT)/obj/lua_host.o: $(HERE)lua_host.cpp $(HOST)/lua_gpu.hpp
	@mkdir -p $(OUT)/obj
	$(REF_CXX) $(CXXFLAGS) -I$(HERE)../include -c $< -o $@

$(OUT)/dort_lua: $(LUA_OBJS) $(GPU_LIB)
	$(REF_CXX) -o $@ $(filter %.o,$^) -Wl,--gc-sections -L$(dir $(GPU_LIB)) -ldortgpu \
	  -Wl,-rpath,'$$ORIGIN/../../dort_amd' -lpthread -lm -ldl

lua: $(OUT)/dort_lua

clean:
	rm -rf $(OUT)

.PHONY: all clean gpu lua
