// compare.hip -- dort's image acceptance rule as a built-in comparator on the GPU.
//
//  * dgpu_test_convergence replaces test_convergence (src/dort/convergence_test.cpp:10-101):
//    for every depth d with (min(w, h) >> d) >= min_tile_size the image is cut into
//    square tiles of side min(w, h) >> d (clipped at the right and bottom edges), and
//    each tile's mean difference (max over channels of tested - reference) must stay
//    within  -stddev / sqrt(pixels) * Phi^-1(p_tile / 2) + bias,  where
//      stddev = max(variation * global reference mean, variation * tile reference mean)
//      bias   = min(bias * global reference mean, bias * tile reference mean)  (max over channels)
//      p_tile = 1 - (1 - p_value)^(1 / number of tiles tested).
//    The first failing tile (smallest depth, then row-major) is reported in dort's words.
//  * dgpu_bias_variance replaces bias_variance (src/dort/lua_image.cpp:179-211).
//  * Phi^-1 is the rational approximation of Abramowitz & Stegun 26.2.23, which dort's
//    normal_cdf_inverse (src/dort/math.cpp:4-19) uses.
// Tile sums are reduced on the GPU in double precision (dort sums pixels in float, one
// after the other; the decisions agree except for tiles within rounding of the threshold).
// One workgroup reduces one tile of one depth; the images stay where they are (host
// buffers are copied to the device once).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dortgpu.h"
#include "dgpu_internal.hpp"
#include "status.hpp"

namespace dgpu {

struct TileGrid {
  uint32_t w, h, ts, tiles_x;
};

// sums[6 * tile + c] = reference channel sum, sums[6 * tile + 3 + c] = tested channel sum
__global__ __launch_bounds__(256) void k_tile_sums(const float* __restrict__ ref, const float* __restrict__ tst,
                                                   TileGrid g, double* __restrict__ sums) {
  const uint32_t tile = blockIdx.x;
  const uint32_t tx = tile % g.tiles_x, ty = tile / g.tiles_x;
  const uint32_t x0 = tx * g.ts, y0 = ty * g.ts;
  const uint32_t x1 = min(x0 + g.ts, g.w), y1 = min(y0 + g.ts, g.h);
  const uint32_t tw = x1 - x0, n = tw * (y1 - y0);
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
    size_t px = (size_t)(y0 + k / tw) * g.w + (x0 + k % tw);
    for (int c = 0; c < 3; ++c) {
      acc[c] += (double)ref[3 * px + c];
      acc[3 + c] += (double)tst[3 * px + c];
    }
  }
  __shared__ double red[6][256];
  for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int c = 0; c < 6; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 6) sums[6 * (size_t)tile + threadIdx.x] = red[threadIdx.x][0];
}

// out[0] = sum of (ref - tested) over the rectangle's channels, out[1] = sum of squares
__global__ __launch_bounds__(256) void k_bias_variance(const float* __restrict__ ref, const float* __restrict__ tst,
                                                       uint32_t w, uint32_t x0, uint32_t y0, uint32_t x1,
                                                       uint32_t y1, double* __restrict__ out) {
  const uint32_t rw = x1 - x0, n = rw * (y1 - y0);
  double b = 0.0, v = 0.0;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    size_t px = (size_t)(y0 + k / rw) * w + (x0 + k % rw);
    for (int c = 0; c < 3; ++c) {
      double d = (double)ref[3 * px + c] - (double)tst[3 * px + c];
      b += d;
      v += d * d;
    }
  }
  __shared__ double rb[256], rv[256];
  rb[threadIdx.x] = b;
  rv[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) { rb[threadIdx.x] += rb[threadIdx.x + s]; rv[threadIdx.x] += rv[threadIdx.x + s]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(out, rb[0]);
    atomicAdd(out + 1, rv[0]);
  }
}

}  // namespace dgpu

namespace {

using namespace dgpu;

// Abramowitz & Stegun 26.2.23: the upper-tail quantile for p in (0, 0.5].
float as_rational(float t) {
  const float c0 = 2.515517f, c1 = 0.802853f, c2 = 0.010328f;
  const float d1 = 1.432788f, d2 = 0.189269f, d3 = 0.001308f;
  return t - ((c2 * t + c1) * t + c0) / (((d3 * t + d2) * t + d1) * t + 1.f);
}

float normal_cdf_inverse(float p) {
  if (p < 0.5f) return -as_rational(std::sqrt(-2.f * std::log(p)));
  return as_rational(std::sqrt(-2.f * std::log(1.f - p)));
}

// Device copies of the two images (or the caller's device pointers).
struct Images {
  const float* ref = nullptr;
  const float* tst = nullptr;
  float* own = nullptr;
  ~Images() { if (own) (void)hipFree(own); }
};

dgpu_status stage(const float* reference, const float* tested, size_t n, uint32_t flags, hipStream_t st,
                  Images& im) {
  if (flags & DGPU_BUF_DEVICE) {
    im.ref = reference;
    im.tst = tested;
    return DGPU_OK;
  }
  if (hipMalloc(&im.own, 2 * n * sizeof(float)) != hipSuccess) return fail(DGPU_ENOMEM, "hipMalloc failed");
  if (hipMemcpyAsync(im.own, reference, n * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(im.own + n, tested, n * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess)
    return fail(DGPU_EDEVICE, "image upload failed");
  im.ref = im.own;
  im.tst = im.own + n;
  return DGPU_OK;
}

std::string f2s(float v) { return std::to_string(v); }

}  // namespace

extern "C" {

dgpu_status dgpu_test_convergence(dgpu_ctx* ctx, const float* reference, const float* tested, uint32_t width,
                                  uint32_t height, int32_t min_tile_size, float variation, float bias,
                                  float p_value, uint32_t flags, int32_t* passed, char* msg, size_t msg_size) {
  if (!ctx || !reference || !tested || !passed) return fail(DGPU_EINVAL, "null argument");
  if (width == 0 || height == 0) return fail(DGPU_EINVAL, "empty image");
  if (min_tile_size <= 0) return fail(DGPU_EINVAL, "min_tile_size must be positive");
  if (!(p_value > 0.f && p_value < 1.f)) return fail(DGPU_EINVAL, "p_value must be in (0, 1)");
  if ((uint32_t)min_tile_size > std::min(width, height)) return fail(DGPU_EINVAL, "image smaller than min_tile_size");
  std::lock_guard<std::mutex> lk(ctx_mutex(ctx));
  if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return fail(DGPU_EDEVICE, "hipSetDevice failed");
  hipStream_t st = ctx_stream(ctx);
  const size_t n = (size_t)width * height * 3;
  Images im;
  dgpu_status s = stage(reference, tested, n, flags, st, im);
  if (s != DGPU_OK) return s;

  const uint32_t m = std::min(width, height);
  uint32_t max_depth = 0;
  while ((m >> (max_depth + 1)) >= (uint32_t)min_tile_size) ++max_depth;
  std::vector<TileGrid> grids;
  std::vector<size_t> first;
  size_t total = 0;
  for (uint32_t d = 0; d <= max_depth; ++d) {
    uint32_t ts = m >> d;
    TileGrid g{width, height, ts, (width + ts - 1) / ts};
    grids.push_back(g);
    first.push_back(total);
    total += (size_t)g.tiles_x * ((height + ts - 1) / ts);
  }
  double* dsums = nullptr;
  if (hipMalloc(&dsums, total * 6 * sizeof(double)) != hipSuccess) return fail(DGPU_ENOMEM, "hipMalloc failed");
  std::vector<double> sums(total * 6);
  hipError_t e = hipSuccess;
  for (size_t d = 0; d < grids.size() && e == hipSuccess; ++d) {
    uint32_t tiles = (uint32_t)(total - first[d]);
    if (d + 1 < grids.size()) tiles = (uint32_t)(first[d + 1] - first[d]);
    hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(256), 0, st, im.ref, im.tst, grids[d], dsums + 6 * first[d]);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(sums.data(), dsums, sums.size() * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(dsums);
  if (e != hipSuccess) return fail(DGPU_EDEVICE, "tile sums failed: %s", hipGetErrorString(e));

  // thresholds as dort computes them, in float
  float test_count = float(((1u << (2 * max_depth + 2)) - 1u) / 3u);
  float test_p_value = 1.f - std::pow(1.f - p_value, 1.f / test_count);
  float inv_phi = normal_cdf_inverse(test_p_value * 0.5f);
  float ref_mean[3];
  for (int c = 0; c < 3; ++c) ref_mean[c] = (float)(sums[c] / ((double)width * height));  // depth 0 is one tile
  auto max3 = [](float a, float b, float c) { return std::max(a, std::max(b, c)); };
  std::string err;
  for (size_t d = 0; d < grids.size() && err.empty(); ++d) {
    const TileGrid& g = grids[d];
    uint32_t tiles_y = (height + g.ts - 1) / g.ts;
    for (uint32_t ty = 0; ty < tiles_y && err.empty(); ++ty) {
      for (uint32_t tx = 0; tx < g.tiles_x; ++tx) {
        const double* sm = &sums[6 * (first[d] + (size_t)ty * g.tiles_x + tx)];
        uint32_t x0 = tx * g.ts, y0 = ty * g.ts;
        uint32_t x1 = std::min(x0 + g.ts, width), y1 = std::min(y0 + g.ts, height);
        double cnt = (double)(x1 - x0) * (double)(y1 - y0);
        float tr[3], tt[3];
        for (int c = 0; c < 3; ++c) { tr[c] = (float)(sm[c] / cnt); tt[c] = (float)(sm[3 + c] / cnt); }
        float mean_diff = max3(tt[0] - tr[0], tt[1] - tr[1], tt[2] - tr[2]);
        float sd = std::max(max3(ref_mean[0] * variation, ref_mean[1] * variation, ref_mean[2] * variation),
                            max3(tr[0] * variation, tr[1] * variation, tr[2] * variation));
        float bi = std::min(max3(ref_mean[0] * bias, ref_mean[1] * bias, ref_mean[2] * bias),
                            max3(tr[0] * bias, tr[1] * bias, tr[2] * bias));
        float max_diff = -sd / std::sqrt((float)cnt) * inv_phi + bi;
        if (std::fabs(mean_diff) > max_diff) {
          err = "tile (" + std::to_string(x0) + "," + std::to_string(y0) + "):(" + std::to_string(x1) + "," +
                std::to_string(y1) + ") has diff " + f2s(mean_diff) + ", but threshold is " + f2s(max_diff);
          break;
        }
      }
    }
  }
  *passed = err.empty() ? 1 : 0;
  if (msg && msg_size) {
    std::strncpy(msg, err.c_str(), msg_size - 1);
    msg[msg_size - 1] = '\0';
  }
  return DGPU_OK;
}

dgpu_status dgpu_bias_variance(dgpu_ctx* ctx, const float* reference, const float* tested, uint32_t width,
                               uint32_t height, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1,
                               uint32_t flags, double* bias_out, double* variance_out) {
  if (!ctx || !reference || !tested || !bias_out || !variance_out) return fail(DGPU_EINVAL, "null argument");
  if (x1 > width || y1 > height) return fail(DGPU_EINVAL, "rect is out of the image");
  *bias_out = 0.0;
  *variance_out = 0.0;
  if (x0 >= x1 || y0 >= y1) return DGPU_OK;
  std::lock_guard<std::mutex> lk(ctx_mutex(ctx));
  if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return fail(DGPU_EDEVICE, "hipSetDevice failed");
  hipStream_t st = ctx_stream(ctx);
  Images im;
  dgpu_status s = stage(reference, tested, (size_t)width * height * 3, flags, st, im);
  if (s != DGPU_OK) return s;
  double* d = nullptr;
  if (hipMalloc(&d, 2 * sizeof(double)) != hipSuccess) return fail(DGPU_ENOMEM, "hipMalloc failed");
  double h[2] = {0.0, 0.0};
  hipError_t e = hipMemcpyAsync(d, h, sizeof h, hipMemcpyHostToDevice, st);
  uint64_t n = (uint64_t)(x1 - x0) * (y1 - y0);
  uint32_t blocks = (uint32_t)std::min<uint64_t>(1024, (n + 255) / 256);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_bias_variance, dim3(blocks), dim3(256), 0, st, im.ref, im.tst, width, x0, y0, x1, y1, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof h, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(DGPU_EDEVICE, "bias_variance failed: %s", hipGetErrorString(e));
  *bias_out = h[0];
  *variance_out = h[1];
  return DGPU_OK;
}

}  // extern "C"
