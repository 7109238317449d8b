// path_order.hpp -- the order in which the paths of one sample visit the pixels.
//
// Paths of a render are numbered sample-major: path g is sample g / npix at position
// g % npix.  Positions visit the pixels in 8x8 tiles: tiles in row-major order, pixels
// row-major inside a tile, the tiles of the right column and bottom row clipped to the
// image.  A wave's 64 consecutive paths then trace one 8x8 block of the image instead
// of a 64-pixel strip of a row, and the queues that shade compacts in path order keep
// that locality.  The random numbers are keyed by pixel and sample, so the order
// changes no result.  Plain integer code, shared by the kernels and host-side checks.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define DG_ORDER_FN __host__ __device__ __forceinline__
#else
#define DG_ORDER_FN inline
#endif

namespace dgpu {

// The pixel (y * xr + x) at position q in [0, xr * yr).
DG_ORDER_FN uint32_t tile_pixel(uint32_t q, uint32_t xr, uint32_t yr) {
  uint32_t band = q / (8u * xr);
  uint32_t r = q - band * 8u * xr;
  uint32_t h = yr - 8u * band < 8u ? yr - 8u * band : 8u;
  uint32_t tile = r / (8u * h);
  uint32_t within = r - tile * 8u * h;
  uint32_t w = xr - 8u * tile < 8u ? xr - 8u * tile : 8u;
  uint32_t y = 8u * band + within / w, x = 8u * tile + within % w;
  return y * xr + x;
}

// The position of pixel px: the inverse of tile_pixel.
DG_ORDER_FN uint32_t tile_pos(uint32_t px, uint32_t xr, uint32_t yr) {
  uint32_t x = px % xr, y = px / xr;
  uint32_t band = y / 8u, tile = x / 8u;
  uint32_t h = yr - 8u * band < 8u ? yr - 8u * band : 8u;
  uint32_t w = xr - 8u * tile < 8u ? xr - 8u * tile : 8u;
  return band * 8u * xr + tile * 8u * h + (y % 8u) * w + x % 8u;
}

// Unsigned 32-bit division by an invariant divisor d >= 1 as a multiply-high and shifts
// (Granlund & Montgomery 1994, "Division by invariant integers using multiplication",
// figure 4.1): with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
// n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0), t = mulhi(m, n), exact for
// every 32-bit n.
struct FastDiv {
  uint32_t d, m, s1, s2;
};

inline FastDiv make_fast_div(uint32_t d) {
  uint32_t l = 0;
  while (l < 32u && (1ull << l) < d) ++l;
  FastDiv f;
  f.d = d;
  f.m = (uint32_t)((((1ull << l) - d) << 32) / d + 1ull);
  f.s1 = l < 1u ? l : 1u;
  f.s2 = l > 0u ? l - 1u : 0u;
  return f;
}

DG_ORDER_FN uint32_t fast_div(uint32_t n, const FastDiv& f) {
  uint32_t t = (uint32_t)(((uint64_t)f.m * n) >> 32);
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

// The pixel order of one image with its divisions precomputed, for the kernels that map
// a path to its pixel at every bounce.
struct PathOrder {
  FastDiv npix;      // xr * yr
  FastDiv band;      // 8 * xr: positions per band of tile rows
  uint32_t xr, yr;
  uint32_t full_bands, full_tiles;  // bands of 8 rows, tiles of 8 columns
};

inline PathOrder make_path_order(uint32_t xr, uint32_t yr) {
  PathOrder o;
  o.npix = make_fast_div(xr * yr);
  o.band = make_fast_div(8u * xr);
  o.xr = xr;
  o.yr = yr;
  o.full_bands = yr / 8u;
  o.full_tiles = xr / 8u;
  return o;
}

// tile_pixel(q) with the divisions of a whole tile as shifts; the clipped tiles of the
// right column and bottom row divide by their width or height (at most 7).
DG_ORDER_FN uint32_t order_pixel(const PathOrder& o, uint32_t q) {
  uint32_t band = fast_div(q, o.band);
  uint32_t r = q - band * 8u * o.xr;
  uint32_t tile, within, w;
  if (band < o.full_bands) {
    tile = r >> 6;
    within = r & 63u;
  } else {
    uint32_t h = o.yr - 8u * band;
    tile = r / (8u * h);
    within = r - tile * 8u * h;
  }
  w = tile < o.full_tiles ? 8u : o.xr - 8u * tile;
  uint32_t y, x;
  if (w == 8u) {
    y = within >> 3;
    x = within & 7u;
  } else {
    y = within / w;
    x = within - y * w;
  }
  return (8u * band + y) * o.xr + 8u * tile + x;
}

// Position t (< 2^32) counted from the start of sample s0: the path's sample and pixel.
DG_ORDER_FN void order_sample_pixel(const PathOrder& o, uint32_t s0, uint32_t t, uint32_t& sample,
                                    uint32_t& pixel) {
  uint32_t ds = fast_div(t, o.npix);
  sample = s0 + ds;
  pixel = order_pixel(o, t - ds * o.npix.d);
}

}  // namespace dgpu
