// io_robustness_check.cpp -- feeds the host readers of libdortgpu.so (dort_amd/csrc/io.cpp:
// dgpu_read_ply, dgpu_read_rgbe, dgpu_read_png) valid files and damaged copies of them,
// for a build with AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitizers.py).
//
// Each valid file is written first (PLY in ascii and both binary byte orders; RGBE flat and
// run-length encoded; PNG through dgpu_write_png, plus a stored-block deflate stream) and
// must read back exactly.  Then every truncation at a set of lengths and a seeded series of
// byte flips is read: a reader may fail with a status, but must not read or write out of
// bounds, overflow a signed integer or leak (the sanitizers abort on any of these).
//
//   g++ -std=c++17 -fsanitize=address,undefined ... io_robustness_check.cpp io.cpp status.cpp
//   ./io_robustness_check DIR      (prints OK and the number of reads)
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../include/dortgpu.h"

namespace {

std::string g_dir;
long g_reads = 0;

std::vector<uint8_t> slurp(const std::string& path) {
  std::vector<uint8_t> b;
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return b;
  int c;
  while ((c = std::fgetc(f)) != EOF) b.push_back((uint8_t)c);
  std::fclose(f);
  return b;
}

void spill(const std::string& path, const std::vector<uint8_t>& b) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) { std::printf("FAIL cannot write %s\n", path.c_str()); std::exit(1); }
  if (!b.empty()) std::fwrite(b.data(), 1, b.size(), f);
  std::fclose(f);
}

[[noreturn]] void fail(const char* what) {
  std::printf("FAIL %s: %s\n", what, dgpu_last_error());
  std::exit(1);
}

// Reads `path` with every reader; the outputs are freed whatever the status.
void read_all(const std::string& path) {
  {
    float* p = nullptr;
    uint32_t* i = nullptr;
    uint32_t nv = 0, nt = 0;
    dgpu_read_ply(path.c_str(), &p, &nv, &i, &nt);
    dgpu_free(p);
    dgpu_free(i);
  }
  {
    float* rgb = nullptr;
    uint32_t w = 0, h = 0;
    dgpu_read_rgbe(path.c_str(), &rgb, &w, &h);
    dgpu_free(rgb);
  }
  {
    uint8_t* rgb = nullptr;
    uint32_t w = 0, h = 0;
    dgpu_read_png(path.c_str(), &rgb, &w, &h);
    dgpu_free(rgb);
  }
  g_reads += 3;
}

// Truncations (every length up to 256, then a spread) and seeded byte flips of `good`.
void damage(const std::vector<uint8_t>& good, uint32_t seed) {
  const std::string path = g_dir + "/damaged.bin";
  std::vector<size_t> cuts;
  for (size_t n = 0; n < good.size() && n < 256; ++n) cuts.push_back(n);
  for (size_t n = 256; n < good.size(); n += 1 + good.size() / 97) cuts.push_back(n);
  for (size_t n : cuts) {
    spill(path, std::vector<uint8_t>(good.begin(), good.begin() + (long)n));
    read_all(path);
  }
  std::mt19937 rng(seed);
  for (int k = 0; k < 400; ++k) {
    std::vector<uint8_t> b = good;
    int flips = 1 + (int)(rng() % 4);
    for (int f = 0; f < flips; ++f) {
      size_t at = rng() % b.size();
      // half the flips land in the first 64 bytes, where the headers are
      if (rng() & 1) at = rng() % std::min<size_t>(64, b.size());
      b[at] ^= (uint8_t)(1u << (rng() % 8));
    }
    spill(path, b);
    read_all(path);
  }
}

std::vector<uint8_t> ply_file(const char* format) {
  // a tetrahedron with an extra property and a quad face (split into two triangles)
  const float v[5][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 1}};
  std::string h = std::string("ply\nformat ") + format +
                  " 1.0\ncomment damaged copies must fail cleanly\nelement vertex 5\n"
                  "property float x\nproperty float y\nproperty float z\nproperty uchar tag\n"
                  "element face 3\nproperty list uchar int vertex_indices\nend_header\n";
  std::vector<uint8_t> b(h.begin(), h.end());
  const int faces[3][4] = {{3, 0, 1, 2}, {3, 0, 1, 3}, {4, 1, 2, 4}};
  const int face_idx[3][4] = {{0, 1, 2, -1}, {0, 1, 3, -1}, {1, 2, 4, 3}};
  const bool ascii = std::strcmp(format, "ascii") == 0;
  const bool big = std::strcmp(format, "binary_big_endian") == 0;
  auto put = [&](const void* p, size_t n) {
    const uint8_t* c = (const uint8_t*)p;
    for (size_t i = 0; i < n; ++i) b.push_back(c[big ? n - 1 - i : i]);
  };
  for (const auto& p : v) {
    if (ascii) {
      char line[96];
      int n = std::snprintf(line, sizeof line, "%g %g %g 7\n", p[0], p[1], p[2]);
      b.insert(b.end(), line, line + n);
    } else {
      for (float x : p) put(&x, 4);
      uint8_t tag = 7;
      put(&tag, 1);
    }
  }
  for (int f = 0; f < 3; ++f) {
    int n = faces[f][0];
    if (ascii) {
      std::string line = std::to_string(n);
      for (int k = 0; k < n; ++k) line += " " + std::to_string(face_idx[f][k]);
      line += "\n";
      b.insert(b.end(), line.begin(), line.end());
    } else {
      uint8_t cnt = (uint8_t)n;
      put(&cnt, 1);
      for (int k = 0; k < n; ++k) {
        int32_t idx = face_idx[f][k];
        put(&idx, 4);
      }
    }
  }
  return b;
}

void check_ply() {
  for (const char* fmt : {"ascii", "binary_little_endian", "binary_big_endian"}) {
    const std::string path = g_dir + "/mesh.ply";
    std::vector<uint8_t> good = ply_file(fmt);
    spill(path, good);
    float* p = nullptr;
    uint32_t* idx = nullptr;
    uint32_t nv = 0, nt = 0;
    if (dgpu_read_ply(path.c_str(), &p, &nv, &idx, &nt) != DGPU_OK) fail("valid PLY");
    if (nv != 5 || nt != 4 || p[3] != 1.f || idx[3 * 3 + 2] != 3) fail("PLY contents");
    dgpu_free(p);
    dgpu_free(idx);
    damage(good, 11);
  }
}

void check_rgbe() {
  // width 5: flat scanlines; width 64: run-length encoded ones
  for (uint32_t w : {5u, 64u}) {
    const uint32_t h = 9;
    std::vector<float> img(size_t(w) * h * 3);
    for (size_t i = 0; i < img.size(); ++i) img[i] = (i % 7 == 0) ? 0.f : float((i * 37) % 101) / 13.f;
    const std::string path = g_dir + "/img.hdr";
    if (dgpu_write_rgbe(path.c_str(), img.data(), w, h) != DGPU_OK) fail("write RGBE");
    float* back = nullptr;
    uint32_t bw = 0, bh = 0;
    if (dgpu_read_rgbe(path.c_str(), &back, &bw, &bh) != DGPU_OK || bw != w || bh != h) fail("valid RGBE");
    dgpu_free(back);
    damage(slurp(path), 12 + w);
  }
}

void check_png() {
  for (uint32_t w : {1u, 7u, 40u}) {
    const uint32_t h = 6;
    std::vector<uint8_t> img(size_t(w) * h * 3);
    for (size_t i = 0; i < img.size(); ++i) img[i] = (uint8_t)((i * 29) ^ (i >> 3));
    const std::string path = g_dir + "/img.png";
    if (dgpu_write_png(path.c_str(), img.data(), w, h) != DGPU_OK) fail("write PNG");
    uint8_t* back = nullptr;
    uint32_t bw = 0, bh = 0;
    if (dgpu_read_png(path.c_str(), &back, &bw, &bh) != DGPU_OK || bw != w || bh != h ||
        std::memcmp(back, img.data(), img.size()) != 0)
      fail("valid PNG");
    dgpu_free(back);
    damage(slurp(path), 20 + w);
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: io_robustness_check DIR\n");
    return 2;
  }
  g_dir = argv[1];
  check_ply();
  check_rgbe();
  check_png();
  std::printf("OK %ld reads\n", g_reads);
  return 0;
}
