// path_order_check.cpp -- CPU check of the kernels' pixel order (dort_amd/csrc/path_order.hpp),
// compiled by tests/test_path_order.py.
//
// For every image size checked: tile_pixel is a bijection of [0, xr*yr) onto the pixels
// and tile_pos is its inverse; every run of positions that fills a whole 8x8 tile covers
// exactly one aligned 8x8 block; in every run of 64 positions the pixels lie within a
// 16x16 box (a wave never spans more than two neighbouring clipped tiles on an edge);
// and the tile bands follow each other top to bottom.  Prints "OK <sizes>" or the first
// failure and exits non-zero.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "path_order.hpp"

using dgpu::tile_pixel;
using dgpu::tile_pos;

static int check(uint32_t xr, uint32_t yr) {
  const uint32_t n = xr * yr;
  std::vector<uint8_t> seen(n, 0);
  for (uint32_t q = 0; q < n; ++q) {
    uint32_t px = tile_pixel(q, xr, yr);
    if (px >= n) { printf("FAIL %ux%u: q %u -> pixel %u out of range\n", xr, yr, q, px); return 1; }
    if (seen[px]++) { printf("FAIL %ux%u: pixel %u reached twice\n", xr, yr, px); return 1; }
    if (tile_pos(px, xr, yr) != q) { printf("FAIL %ux%u: tile_pos(tile_pixel(%u)) = %u\n", xr, yr, q, tile_pos(px, xr, yr)); return 1; }
  }
  // whole tiles: 64 consecutive positions starting at the tile's first position
  for (uint32_t ty = 0; ty + 8 <= yr; ty += 8)
    for (uint32_t tx = 0; tx + 8 <= xr; tx += 8) {
      uint32_t q0 = tile_pos(ty * xr + tx, xr, yr);
      for (uint32_t k = 0; k < 64; ++k) {
        uint32_t px = tile_pixel(q0 + k, xr, yr), x = px % xr, y = px / xr;
        if (x < tx || x >= tx + 8 || y < ty || y >= ty + 8 || x != tx + k % 8 || y != ty + k / 8) {
          printf("FAIL %ux%u: tile (%u,%u) position %u is pixel (%u,%u)\n", xr, yr, tx, ty, k, x, y);
          return 1;
        }
      }
    }
  // any 64 consecutive positions aligned to 64 stay within a small box, except where a
  // band of clipped tiles is thinner than 8 rows or columns (edges of the image)
  if (xr % 8 == 0 && yr % 8 == 0)
    for (uint32_t q0 = 0; q0 + 64 <= n; q0 += 64) {
      uint32_t x0 = ~0u, x1 = 0, y0 = ~0u, y1 = 0;
      for (uint32_t k = 0; k < 64; ++k) {
        uint32_t px = tile_pixel(q0 + k, xr, yr), x = px % xr, y = px / xr;
        x0 = x < x0 ? x : x0; x1 = x > x1 ? x : x1; y0 = y < y0 ? y : y0; y1 = y > y1 ? y : y1;
      }
      if (x1 - x0 > 7 || y1 - y0 > 7) { printf("FAIL %ux%u: wave at %u spans %ux%u\n", xr, yr, q0, x1 - x0 + 1, y1 - y0 + 1); return 1; }
    }
  // bands in order: the row of tile_pixel(q) never decreases by more than 7
  uint32_t maxy = 0;
  for (uint32_t q = 0; q < n; ++q) {
    uint32_t y = tile_pixel(q, xr, yr) / xr;
    if (y + 7 < maxy) { printf("FAIL %ux%u: position %u goes back to row %u after row %u\n", xr, yr, q, y, maxy); return 1; }
    maxy = y > maxy ? y : maxy;
  }
  return 0;
}

int main() {
  int sizes = 0;
  for (uint32_t xr = 1; xr <= 40; ++xr)
    for (uint32_t yr = 1; yr <= 40; ++yr, ++sizes)
      if (check(xr, yr)) return 1;
  const uint32_t big[][2] = {{1024, 1024}, {256, 256}, {512, 512}, {768, 768}, {1023, 769},
                             {4096, 3}, {3, 4096}, {300, 217}, {64, 48}, {4096, 2048}};
  for (auto& s : big) { if (check(s[0], s[1])) return 1; ++sizes; }
  printf("OK %d\n", sizes);
  return 0;
}
