// ref_io_driver.cpp -- a C ABI over dort's file readers/writers and its convergence
// test (TEST INFRASTRUCTURE ONLY).
//
// Compiled with dort's sources where they lie (oracle/ref.mk) into
// oracle/_ref/libdortref.so, so the tests can compare the library's PLY / RGBE code and
// its GPU comparator with dort's own:
//   ref_read_ply          read_ply_to_ply_mesh (src/dort/ply_mesh.cpp, over rply)
//   ref_read_hdr          read_image_f (src/dort/image.cpp, stb_image)
//   ref_write_hdr         write_image_rgbe (src/dort/image.cpp, stb_image_write)
//   ref_read_png          read_image_8 (src/dort/image.cpp, stb_image; also JPEG / BMP)
//   ref_write_png         write_image_png (src/dort/image.cpp, stb_image_write)
//   ref_test_convergence  test_convergence (src/dort/convergence_test.cpp), one thread
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include <stb_image.h>

#include "dort/convergence_test.hpp"
#include "dort/ctx.hpp"
#include "dort/geometry.hpp"
#include "dort/image.hpp"
#include "dort/ply_mesh.hpp"
#include "dort/thread_pool.hpp"

using namespace dort;

extern "C" {

int ref_read_ply(const char* path, float** pos, uint32_t* nv, uint32_t** ind, uint32_t* nt) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return 1;
  PlyMesh m;
  m.triangle_count = 0;
  bool ok = read_ply_to_ply_mesh(f, m);
  std::fclose(f);
  if (!ok) return 2;
  *nv = (uint32_t)m.points.size();
  *nt = (uint32_t)(m.vertices.size() / 3);
  *pos = (float*)std::malloc(3 * m.points.size() * sizeof(float) + 4);
  *ind = (uint32_t*)std::malloc(m.vertices.size() * sizeof(uint32_t) + 4);
  for (size_t i = 0; i < m.points.size(); ++i)
    for (int k = 0; k < 3; ++k) (*pos)[3 * i + k] = m.points[i].v[k];
  if (!m.vertices.empty()) std::memcpy(*ind, m.vertices.data(), m.vertices.size() * sizeof(uint32_t));
  return 0;
}

void ref_free(void* p) { std::free(p); }

int ref_read_hdr(const char* path, float** rgb, uint32_t* w, uint32_t* h) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return 1;
  Image<PixelRgbFloat> img = read_image_f(f);
  std::fclose(f);
  *w = (uint32_t)img.res.x;
  *h = (uint32_t)img.res.y;
  *rgb = (float*)std::malloc(img.storage.size() * sizeof(float) + 4);
  std::memcpy(*rgb, img.storage.data(), img.storage.size() * sizeof(float));
  return 0;
}

int ref_write_hdr(const char* path, const float* rgb, uint32_t w, uint32_t h) {
  Image<PixelRgbFloat> img(w, h);
  std::memcpy(img.storage.data(), rgb, img.storage.size() * sizeof(float));
  FILE* f = std::fopen(path, "wb");
  if (!f) return 1;
  write_image_rgbe(f, img);
  std::fclose(f);
  return 0;
}

int ref_read_png(const char* path, uint8_t** rgb, uint32_t* w, uint32_t* h) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return 1;
  // read_image_8 asserts on a file stb cannot decode; probe it first
  int x = 0, y = 0, n = 0;
  bool ok = stbi_info_from_file(f, &x, &y, &n) != 0;
  if (!ok) { std::fclose(f); return 2; }
  Image<PixelRgb8> img = read_image_8(f);
  std::fclose(f);
  *w = (uint32_t)img.res.x;
  *h = (uint32_t)img.res.y;
  *rgb = (uint8_t*)std::malloc(img.storage.size() + 4);
  std::memcpy(*rgb, img.storage.data(), img.storage.size());
  return 0;
}

int ref_write_png(const char* path, const uint8_t* rgb, uint32_t w, uint32_t h) {
  Image<PixelRgb8> img(w, h);
  std::memcpy(img.storage.data(), rgb, img.storage.size());
  FILE* f = std::fopen(path, "wb");
  if (!f) return 1;
  write_image_png(f, img);
  std::fclose(f);
  return 0;
}

int ref_test_convergence(const float* ref, const float* tested, uint32_t w, uint32_t h, int32_t min_tile_size,
                         float variation, float bias, float p_value, char* msg, size_t msg_size) {
  Image<PixelRgbFloat> a(w, h), b(w, h);
  std::memcpy(a.storage.data(), ref, a.storage.size() * sizeof(float));
  std::memcpy(b.storage.data(), tested, b.storage.size() * sizeof(float));
  CtxG ctx;
  ctx.pool = std::make_shared<ThreadPool>();
  ctx.pool->start(1);
  std::string err = test_convergence(ctx, a, b, min_tile_size, variation, bias, p_value);
  ctx.pool->stop();
  std::strncpy(msg, err.c_str(), msg_size - 1);
  msg[msg_size - 1] = '\0';
  return err.empty() ? 1 : 0;
}

}  // extern "C"
