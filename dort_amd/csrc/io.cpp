// io.cpp -- scene and image files for the drop-in: PLY meshes and Radiance RGBE images.
//
// Replaces dort's readers/writers at the two ends of the pt path:
//  * read_ply_callbacks / read_ply_to_mesh (src/dort/ply_mesh.cpp:9-103, over rply):
//    vertex x, y, z and the faces' vertex_indices lists; a face (i0, i1, ..., ik) becomes
//    the fan (i0, i1, i2), (i0, i2, i3), ... in list order, as the rply callback there
//    emits it (:51-65).  Other elements and properties are skipped.
//  * read_image_f / write_image_rgbe (src/dort/image.cpp:28-67, over stb_image):
//    Radiance RGBE with "-Y H +X W" scanlines, flat or run-length encoded, decoded as
//    m * 2^(e - 136) per channel (0 when e == 0) and encoded with Ward's float-to-RGBE
//    rule (mantissa of the largest channel scaled to 256).  Like stb, a file whose first
//    scanline is not run-length encoded is read as flat pixels; unlike stb, a run that
//    would overflow its scanline is an error, and "#?RGBE" headers are accepted too.
// Written from the PLY format description (Turk, 1994) and the Radiance picture format
// (Ward, "Real pixels", Graphics Gems II, 1991).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "status.hpp"

namespace {

using dgpu::fail;

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() { if (f) std::fclose(f); }
};

// ----------------------------------------------------------------- PLY

enum PlyType { PT_NONE, PT_I8, PT_U8, PT_I16, PT_U16, PT_I32, PT_U32, PT_F32, PT_F64 };

PlyType ply_type(const std::string& s) {
  if (s == "char" || s == "int8") return PT_I8;
  if (s == "uchar" || s == "uint8") return PT_U8;
  if (s == "short" || s == "int16") return PT_I16;
  if (s == "ushort" || s == "uint16") return PT_U16;
  if (s == "int" || s == "int32") return PT_I32;
  if (s == "uint" || s == "uint32") return PT_U32;
  if (s == "float" || s == "float32") return PT_F32;
  if (s == "double" || s == "float64") return PT_F64;
  return PT_NONE;
}

size_t ply_size(PlyType t) {
  switch (t) {
    case PT_I8: case PT_U8: return 1;
    case PT_I16: case PT_U16: return 2;
    case PT_I32: case PT_U32: case PT_F32: return 4;
    case PT_F64: return 8;
    default: return 0;
  }
}

struct PlyProp {
  std::string name;
  PlyType type = PT_NONE;       // scalar type, or the list's value type
  PlyType count_type = PT_NONE; // list length type (PT_NONE: scalar)
};

struct PlyElem {
  std::string name;
  uint64_t count = 0;
  std::vector<PlyProp> props;
};

enum PlyFormat { PF_ASCII, PF_LE, PF_BE };

// Reads one value of type t (binary: raw bytes in the file's byte order; ascii: the
// next whitespace-separated token).  Returns false at end of file or on a bad token.
bool ply_value(FILE* f, PlyFormat fmt, PlyType t, double& out) {
  if (fmt == PF_ASCII) {
    char tok[64];
    if (std::fscanf(f, "%63s", tok) != 1) return false;
    char* end = nullptr;
    out = std::strtod(tok, &end);
    return end && *end == '\0';
  }
  unsigned char b[8];
  size_t n = ply_size(t);
  if (std::fread(b, 1, n, f) != n) return false;
  bool host_le = true;
  {
    uint16_t one = 1;
    unsigned char c;
    std::memcpy(&c, &one, 1);
    host_le = c == 1;
  }
  if ((fmt == PF_LE) != host_le) std::reverse(b, b + n);
  switch (t) {
    case PT_I8: { int8_t v; std::memcpy(&v, b, 1); out = v; break; }
    case PT_U8: { uint8_t v; std::memcpy(&v, b, 1); out = v; break; }
    case PT_I16: { int16_t v; std::memcpy(&v, b, 2); out = v; break; }
    case PT_U16: { uint16_t v; std::memcpy(&v, b, 2); out = v; break; }
    case PT_I32: { int32_t v; std::memcpy(&v, b, 4); out = v; break; }
    case PT_U32: { uint32_t v; std::memcpy(&v, b, 4); out = v; break; }
    case PT_F32: { float v; std::memcpy(&v, b, 4); out = v; break; }
    case PT_F64: { double v; std::memcpy(&v, b, 8); out = v; break; }
    default: return false;
  }
  return true;
}

// Header lines up to end_header; false (with msg) on a malformed header.
bool ply_header(FILE* f, PlyFormat& fmt, std::vector<PlyElem>& elems, std::string& msg) {
  char line[1024];
  if (!std::fgets(line, sizeof line, f) || std::strncmp(line, "ply", 3) != 0) {
    msg = "not a PLY file";
    return false;
  }
  bool have_format = false;
  while (std::fgets(line, sizeof line, f)) {
    char w0[64] = "", w1[64] = "", w2[64] = "", w3[64] = "", w4[64] = "";
    int n = std::sscanf(line, "%63s %63s %63s %63s %63s", w0, w1, w2, w3, w4);
    if (n <= 0) continue;
    std::string k = w0;
    if (k == "end_header") {
      if (!have_format) { msg = "PLY header has no format line"; return false; }
      return true;
    }
    if (k == "comment" || k == "obj_info") continue;
    if (k == "format") {
      std::string v = w1;
      if (v == "ascii") fmt = PF_ASCII;
      else if (v == "binary_little_endian") fmt = PF_LE;
      else if (v == "binary_big_endian") fmt = PF_BE;
      else { msg = "unknown PLY format " + v; return false; }
      have_format = true;
    } else if (k == "element") {
      if (n < 3) { msg = "bad PLY element line"; return false; }
      PlyElem e;
      e.name = w1;
      e.count = std::strtoull(w2, nullptr, 10);
      elems.push_back(e);
    } else if (k == "property") {
      if (elems.empty()) { msg = "PLY property before any element"; return false; }
      PlyProp p;
      if (std::string(w1) == "list") {
        if (n < 5) { msg = "bad PLY list property"; return false; }
        p.count_type = ply_type(w2);
        p.type = ply_type(w3);
        p.name = w4;
        if (p.count_type == PT_NONE || p.count_type == PT_F32 || p.count_type == PT_F64 || p.type == PT_NONE) {
          msg = "bad PLY list types";
          return false;
        }
      } else {
        if (n < 3) { msg = "bad PLY property line"; return false; }
        p.type = ply_type(w1);
        p.name = w2;
        if (p.type == PT_NONE) { msg = std::string("unknown PLY type ") + w1; return false; }
      }
      elems.back().props.push_back(p);
    } else {
      msg = "unknown PLY header line: " + k;
      return false;
    }
  }
  msg = "PLY header without end_header";
  return false;
}

}  // namespace

extern "C" {

dgpu_status dgpu_read_ply(const char* path, float** positions, uint32_t* num_vertices, uint32_t** indices,
                          uint32_t* num_triangles) {
  if (!path || !positions || !num_vertices || !indices || !num_triangles)
    return fail(DGPU_EINVAL, "null argument");
  *positions = nullptr;
  *indices = nullptr;
  *num_vertices = 0;
  *num_triangles = 0;
  File file(path, "rb");
  if (!file.f) return fail(DGPU_EINVAL, "cannot open %s", path);
  PlyFormat fmt = PF_ASCII;
  std::vector<PlyElem> elems;
  std::string msg;
  if (!ply_header(file.f, fmt, elems, msg)) return fail(DGPU_EINVAL, "%s: %s", path, msg.c_str());
  std::vector<float> pos;
  std::vector<uint32_t> tri;
  uint64_t nverts = 0;
  for (const PlyElem& e : elems) {
    const bool is_vertex = e.name == "vertex", is_face = e.name == "face";
    int px = -1, py = -1, pz = -1, pf = -1;
    for (size_t k = 0; k < e.props.size(); ++k) {
      const PlyProp& p = e.props[k];
      if (is_vertex && p.count_type == PT_NONE) {
        if (p.name == "x") px = (int)k;
        if (p.name == "y") py = (int)k;
        if (p.name == "z") pz = (int)k;
      }
      if (is_face && p.count_type != PT_NONE && p.name == "vertex_indices") pf = (int)k;
    }
    if (is_vertex) {
      if (e.count > 0xFFFFFFFFull / 3) return fail(DGPU_EINVAL, "%s: too many vertices", path);
      nverts = e.count;
      pos.assign(3 * e.count, 0.f);
    }
    std::vector<double> list;
    for (uint64_t r = 0; r < e.count; ++r) {
      for (size_t k = 0; k < e.props.size(); ++k) {
        const PlyProp& p = e.props[k];
        if (p.count_type == PT_NONE) {
          double v;
          if (!ply_value(file.f, fmt, p.type, v)) return fail(DGPU_EINVAL, "%s: truncated %s data", path, e.name.c_str());
          if (is_vertex) {
            if ((int)k == px) pos[3 * r] = (float)v;
            if ((int)k == py) pos[3 * r + 1] = (float)v;
            if ((int)k == pz) pos[3 * r + 2] = (float)v;
          }
        } else {
          double c;
          if (!ply_value(file.f, fmt, p.count_type, c) || c < 0.0)
            return fail(DGPU_EINVAL, "%s: truncated %s data", path, e.name.c_str());
          list.resize((size_t)c);
          for (size_t j = 0; j < list.size(); ++j)
            if (!ply_value(file.f, fmt, p.type, list[j]))
              return fail(DGPU_EINVAL, "%s: truncated %s data", path, e.name.c_str());
          if ((int)k == pf) {
            // fan: (i0, i1, i2), (i0, i2, i3), ...
            for (size_t j = 2; j < list.size(); ++j) {
              uint32_t a = (uint32_t)list[0], b = (uint32_t)list[j - 1], cidx = (uint32_t)list[j];
              if (list[0] < 0.0 || list[j - 1] < 0.0 || list[j] < 0.0 || a >= nverts || b >= nverts ||
                  cidx >= nverts)
                return fail(DGPU_EINVAL, "%s: face %llu has a vertex index out of range", path,
                            (unsigned long long)r);
              tri.push_back(a);
              tri.push_back(b);
              tri.push_back(cidx);
            }
          }
        }
      }
    }
  }
  if (tri.size() / 3 > 0xFFFFFFFFull) return fail(DGPU_EINVAL, "%s: too many triangles", path);
  float* p = (float*)std::malloc(std::max<size_t>(1, pos.size()) * sizeof(float));
  uint32_t* t = (uint32_t*)std::malloc(std::max<size_t>(1, tri.size()) * sizeof(uint32_t));
  if (!p || !t) {
    std::free(p);
    std::free(t);
    return fail(DGPU_ENOMEM, "out of host memory");
  }
  if (!pos.empty()) std::memcpy(p, pos.data(), pos.size() * sizeof(float));
  if (!tri.empty()) std::memcpy(t, tri.data(), tri.size() * sizeof(uint32_t));
  *positions = p;
  *indices = t;
  *num_vertices = (uint32_t)nverts;
  *num_triangles = (uint32_t)(tri.size() / 3);
  return DGPU_OK;
}

void dgpu_free(void* p) { std::free(p); }

// ----------------------------------------------------------------- RGBE

dgpu_status dgpu_read_rgbe(const char* path, float** rgb, uint32_t* width, uint32_t* height) {
  if (!path || !rgb || !width || !height) return fail(DGPU_EINVAL, "null argument");
  *rgb = nullptr;
  *width = 0;
  *height = 0;
  File file(path, "rb");
  if (!file.f) return fail(DGPU_EINVAL, "cannot open %s", path);
  FILE* f = file.f;
  char line[1024];
  if (!std::fgets(line, sizeof line, f) ||
      (std::strncmp(line, "#?RADIANCE", 10) != 0 && std::strncmp(line, "#?RGBE", 6) != 0))
    return fail(DGPU_EINVAL, "%s: not a Radiance RGBE file", path);
  bool rle_format = false;
  for (;;) {
    if (!std::fgets(line, sizeof line, f)) return fail(DGPU_EINVAL, "%s: truncated header", path);
    if (line[0] == '\n' || line[0] == '\r' || line[0] == '\0') break;
    if (std::strncmp(line, "FORMAT=32-bit_rle_rgbe", 22) == 0) rle_format = true;
  }
  if (!rle_format) return fail(DGPU_EUNSUPPORTED, "%s: FORMAT is not 32-bit_rle_rgbe", path);
  int h = 0, w = 0;
  if (!std::fgets(line, sizeof line, f) || std::sscanf(line, "-Y %d +X %d", &h, &w) != 2 || h <= 0 || w <= 0)
    return fail(DGPU_EUNSUPPORTED, "%s: only \"-Y H +X W\" orientation is supported", path);
  std::vector<float> out((size_t)w * h * 3);
  std::vector<unsigned char> scan((size_t)w * 4);
  auto decode = [&](const unsigned char* e, float* o) {
    if (e[3] == 0) { o[0] = o[1] = o[2] = 0.f; return; }
    float s = std::ldexp(1.0f, (int)e[3] - 136);
    o[0] = e[0] * s; o[1] = e[1] * s; o[2] = e[2] * s;
  };
  for (int y = 0; y < h; ++y) {
    unsigned char hd[4];
    if (std::fread(hd, 1, 4, f) != 4) return fail(DGPU_EINVAL, "%s: truncated pixels", path);
    bool rle = w >= 8 && w < 32768 && hd[0] == 2 && hd[1] == 2 && !(hd[2] & 0x80);
    if (rle && ((int)hd[2] << 8 | hd[3]) != w) return fail(DGPU_EINVAL, "%s: bad scanline length", path);
    if (!rle) {
      // flat pixels (the rest of the image, from here on)
      std::vector<unsigned char> flat((size_t)w * (h - y) * 4);
      std::memcpy(flat.data(), hd, 4);
      if (std::fread(flat.data() + 4, 1, flat.size() - 4, f) != flat.size() - 4)
        return fail(DGPU_EINVAL, "%s: truncated pixels", path);
      for (size_t k = 0; k < (size_t)w * (h - y); ++k) decode(&flat[4 * k], &out[3 * ((size_t)y * w + k)]);
      break;
    }
    for (int c = 0; c < 4; ++c) {
      int x = 0;
      while (x < w) {
        int count = std::fgetc(f);
        if (count == EOF) return fail(DGPU_EINVAL, "%s: truncated run", path);
        if (count > 128) {
          count -= 128;
          int v = std::fgetc(f);
          if (v == EOF || x + count > w) return fail(DGPU_EINVAL, "%s: bad run", path);
          for (int k = 0; k < count; ++k) scan[4 * (x++) + c] = (unsigned char)v;
        } else {
          if (x + count > w) return fail(DGPU_EINVAL, "%s: bad run", path);
          for (int k = 0; k < count; ++k) {
            int v = std::fgetc(f);
            if (v == EOF) return fail(DGPU_EINVAL, "%s: truncated run", path);
            scan[4 * (x++) + c] = (unsigned char)v;
          }
        }
      }
    }
    for (int x = 0; x < w; ++x) decode(&scan[4 * x], &out[3 * ((size_t)y * w + x)]);
  }
  float* p = (float*)std::malloc(out.size() * sizeof(float));
  if (!p) return fail(DGPU_ENOMEM, "out of host memory");
  std::memcpy(p, out.data(), out.size() * sizeof(float));
  *rgb = p;
  *width = (uint32_t)w;
  *height = (uint32_t)h;
  return DGPU_OK;
}

dgpu_status dgpu_write_rgbe(const char* path, const float* rgb, uint32_t width, uint32_t height) {
  if (!path || !rgb) return fail(DGPU_EINVAL, "null argument");
  if (width == 0 || height == 0 || width > 0x7FFFFFFFu || height > 0x7FFFFFFFu)
    return fail(DGPU_EINVAL, "bad image size");
  File file(path, "wb");
  if (!file.f) return fail(DGPU_EINVAL, "cannot open %s for writing", path);
  FILE* f = file.f;
  std::fprintf(f, "#?RADIANCE\n# dort_amd\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", height, width);
  std::vector<unsigned char> px((size_t)width * 4);
  std::vector<unsigned char> enc;
  for (uint32_t y = 0; y < height; ++y) {
    for (uint32_t x = 0; x < width; ++x) {
      const float* c = rgb + 3 * ((size_t)y * width + x);
      float m = std::max(c[0], std::max(c[1], c[2]));
      unsigned char* e = &px[4 * x];
      if (!(m >= 1e-32f)) {
        e[0] = e[1] = e[2] = e[3] = 0;
      } else {
        int ex;
        float s = std::frexp(m, &ex) * 256.0f / m;
        e[0] = (unsigned char)(c[0] * s);
        e[1] = (unsigned char)(c[1] * s);
        e[2] = (unsigned char)(c[2] * s);
        e[3] = (unsigned char)(ex + 128);
      }
    }
    if (width < 8 || width >= 32768) {
      if (std::fwrite(px.data(), 1, px.size(), f) != px.size()) return fail(DGPU_EINVAL, "write failed");
      continue;
    }
    // run-length encoded scanline: each channel as runs (count > 128: repeat) and literals
    enc.clear();
    enc.push_back(2); enc.push_back(2);
    enc.push_back((unsigned char)(width >> 8)); enc.push_back((unsigned char)(width & 0xFF));
    for (int c = 0; c < 4; ++c) {
      uint32_t x = 0;
      while (x < width) {
        uint32_t run = 1;
        while (x + run < width && run < 127 && px[4 * (x + run) + c] == px[4 * x + c]) ++run;
        if (run >= 3) {
          enc.push_back((unsigned char)(128 + run));
          enc.push_back(px[4 * x + c]);
          x += run;
          continue;
        }
        // literal until the next run of 3 or 128 bytes
        uint32_t start = x, n = 0;
        while (x < width && n < 128) {
          if (x + 2 < width && px[4 * x + c] == px[4 * (x + 1) + c] && px[4 * x + c] == px[4 * (x + 2) + c]) break;
          ++x;
          ++n;
        }
        enc.push_back((unsigned char)n);
        for (uint32_t k = 0; k < n; ++k) enc.push_back(px[4 * (start + k) + c]);
      }
    }
    if (std::fwrite(enc.data(), 1, enc.size(), f) != enc.size()) return fail(DGPU_EINVAL, "write failed");
  }
  return DGPU_OK;
}

}  // extern "C"
