// dgpu_internal.hpp -- launch interface between the C-ABI (api.hip) and the kernels.
#pragma once
#include <mutex>

#include "dgpu_device.hpp"

namespace dgpu {

// Device counters of one wavefront chunk (zeroed by raygen).
constexpr uint32_t MAX_BOUNCES = 64;
constexpr uint32_t CNT_QUEUE = 0;        // [b]: extension rays of bounce b
constexpr uint32_t CNT_SHADOW = 64;      // [b]: shadow rays of bounce b
constexpr uint32_t CNT_FETCH_EXT = 128;  // persistent-kernel work counters
constexpr uint32_t CNT_FETCH_SHD = 192;
constexpr uint32_t NUM_COUNTERS = 256;
// Render statistics on the device: paths, extension rays, shadow rays, dropped samples,
// then 16 per-bounce extension and 16 per-bounce shadow ray counts.
constexpr uint32_t STATS_WORDS = 4 + 2 * 16;
constexpr uint32_t MAX_SLOT_LIGHTS = 8;  // sample_all_lights supports up to this many lights
#ifndef DGPU_SHADE_BLOCK
#define DGPU_SHADE_BLOCK 256
#endif
constexpr uint32_t SHADE_BLOCK = DGPU_SHADE_BLOCK;  // threads per shade workgroup (one queue atomic each)
#ifndef DGPU_TRACE_BATCH
#define DGPU_TRACE_BATCH 4
#endif
constexpr uint32_t TRACE_BATCH = DGPU_TRACE_BATCH;  // 64-ray batches a trace wave takes per atomic
// Trace waves refill lanes whose ray has ended from their own batch once this many lanes
// are idle (0: each 64-ray part of the batch is traced to its end before the next):
// any-hit (shadow) rays, and closest-hit (extension) rays after the first bounce.
#ifndef DGPU_REFILL_ANY
#define DGPU_REFILL_ANY 16
#endif
#ifndef DGPU_REFILL_LATE
#define DGPU_REFILL_LATE 32
#endif
// Closest-hit rays of dgpu_trace_closest (thresholds 8, 16 and 24 measured alike on the
// 16M-ray intersection workload: 10.0-10.1 Grays/s against 8.8-8.9 without refills).
#ifndef DGPU_REFILL_API
#define DGPU_REFILL_API 16
#endif
#ifndef DGPU_TRACE_BATCH_ANY
#define DGPU_TRACE_BATCH_ANY 8
#endif
constexpr uint32_t TRACE_BATCH_ANY = DGPU_TRACE_BATCH_ANY;

constexpr uint32_t MAX_BATCH = 255;      // scenes per batch
constexpr uint32_t MAX_CHUNK_SEGS = 64;  // scenes overlapping one chunk of a batch

struct KernelCfg {
  uint32_t width;  // BVH width
  uint32_t lds;    // traversal stack entries in LDS per lane
  uint32_t block;  // traversal workgroup size
  uint32_t grid;   // persistent grid
};

// One scene of a batch render.
struct BatchEntry {
  DevScene S;
  uint32_t xr, yr, npix, sample_begin;
  uint32_t k0, k1, pad0, pad1;
  PathOrder ord;   // the scene's pixel order
  float* film_rgb;
  float* film_w;
};

// The part of one scene's paths that a batch chunk holds: scene-local path indices
// [local_begin, local_end) at chunk slots [slot_begin, ...); film pairs (one per
// distinct pixel) start at pair_begin.  The segment's first path is at position
// q_begin (< npix) of sample s_begin.
struct ChunkSeg {
  uint32_t scene, slot_begin, pair_begin, n_pairs;
  uint64_t local_begin, local_end;
  uint32_t q_begin, s_begin;
};

struct ChunkDesc {
  uint32_t n_chunk, n_segs, n_pairs, pad;
  ChunkSeg segs[MAX_CHUNK_SEGS];
};

struct WaveParams {
  DevScene S;
  uint32_t xr, yr, npix;
  uint32_t min_depth, max_depth, rr_depth;
  uint32_t only_direct, sample_all, strategy;
  uint32_t slots;          // pending NEE slots per path (2 per light considered)
  uint32_t rough;          // some scene of the render has a rough dielectric (shade with that lobe)
  uint32_t k0, k1;         // RNG key
  uint32_t sample_begin;
  uint32_t n_chunk;        // paths in this chunk
  uint64_t g0;             // first global path index of the chunk
  PathOrder ord;           // pixel order of the image
  uint32_t q0, s0;         // the chunk's first path: position q0 (< npix) of sample s0;
                           // chunks are cut so that q0 + n_chunk < 2^32
  float4* thr;             // throughput rgb + flags
  float4* rad;             // radiance accumulated so far
  float4* pend;            // pending NEE contributions [path * slots + k]; w = 1 while visible
  float4* pend_b;          // nee_lag 2: the odd bounces' set (finalize reads both)
  uint32_t nee_lag;        // shade adds a bounce's visible NEE nee_lag bounces later (0 = 1)
  float4* qo[2];           // extension queues: origin + tmin
  float4* qd[2];           //                   direction + tmax
  uint32_t* qpath[2];      //                   path slot
  float4* hit;             // t, b1, b2, leaf-order triangle (bits)
  float4* so;              // shadow queue: origin + tmin
  float4* sd;              //               direction + tmax
  uint2* smeta;            //               pending slot, ignored triangle
  uint32_t* counters;
  unsigned long long* stats;  // paths, extension rays, shadow rays, dropped samples
  uint32_t* spill;
  uint32_t stack_cap;
  float* film_rgb;
  float* film_w;
  // batch rendering
  const BatchEntry* batch;
  uint32_t num_scenes;
  const ChunkDesc* chunks;
  uint32_t* chunk_cur;     // index of the chunk being rendered (advanced on the device)
  // light tracing (dgpu_render_light)
  uint32_t min_len, max_len;  // path lengths in vertices (LightRenderer min_length / max_length)
  float2* fpos;            // film position of each camera-connection shadow ray
  // adjoint replay (dgpu_adjoint)
  float4* adj_L;           // per path: the first pass's radiance; w = 1 if the sample counts
  float4* adj_pend;        // per NEE slot: gradient payload of the pending shadow ray
  float4* adj_cont;        // per path: gradient payload of the last continuation
  const float* adj_dldi;   // dL/dI per pixel (rgb)
  float* adj_g_tex;        // texture-0 gradient, texels * 3 (or null)
  float* adj_g_rough;      // per-material roughness gradient (or null)
  float* adj_image;        // += per-pixel mean radiance of the first pass (or null)
  float adj_inv_spp;
};

struct AdjArgs {
  const float* dldi;  // npix * 3, dL/dI
  float* g_tex;       // texture-0 gradient, texels * 3 (or null)
  float* g_rough;     // per-material roughness gradient (or null)
  float* image;       // optional: += per-pixel mean of the pass-1 path radiance
  float inv_spp;
};

// Paths per adjoint launch (bounds the traversal spill area: one slice per path).
constexpr uint32_t ADJ_CHUNK = 1u << 20;

// Bidirectional path tracing (bdpt.hip): the subpaths' vertex arrays of one chunk.
struct BdptArgsHost {
  float4* lv;        // light subpaths: nv vertices x 4 float4 x n paths
  float4* cv;        // camera subpaths, likewise
  uint32_t n;        // paths per chunk (the arrays' stride)
  uint32_t nv;       // vertex slots per subpath (max_depth + 2)
  uint32_t use_t1;
  float* splat;      // splat sums, npix * 3
  // wavefront form (launch_bdpt_wave): per path
  float4* w0;        // walk state, light walks at [i], camera walks at [n + i]: throughput
                     // factor (light: last f; camera: the pending ray's scale) and the
                     // forward direction pdf
  uint4* w1;         // walk state, likewise: vertices so far, light index, distant light
  uint32_t* nl;      // light subpath vertices | the path's light << 8
  uint32_t* nc;      // camera subpath vertices
  // strategy stage, over a part of m paths of the chunk: slot k of path j at [k * m + j]
  float4* rec;       // weighted contribution, w = 1 (sum) / 2 (splat) / 0
  float2* rec_fp;    // film position of a splat record
  uint32_t ns;       // strategy slots per path: nv (nv + 1) / 2
};
constexpr uint32_t BDPT_VERTEX_BYTES = 64;

bool variant_supported(uint32_t width, uint32_t lds, uint32_t block);
// which: 0 closest-hit API, 1 any-hit API, 2 extend, 3 connect, 4/5 batched extend/connect
// (the bounded API variants share the occupancy of 0/1)
const void* trace_kernel_ptr(int which, uint32_t width, uint32_t lds, uint32_t block);
const void* shade_kernel_ptr(bool batch, bool all, bool rough, bool adj = false);

// bounded: the ray origins lie within the range the box padding covers (fast box test)
hipError_t launch_trace_api(const DevScene& S, const KernelCfg& cfg, const dgpu_rays& rays,
                            const dgpu_hits* hits, uint32_t* occluded, uint32_t n, uint32_t* fetch,
                            uint32_t* spill, bool bounded, hipStream_t st);
hipError_t launch_raygen(const WaveParams& P, uint32_t capacity, bool batch, hipStream_t st);
// k_bdpt over the chunk (P.n_chunk paths) and the film update of its t >= 2 strategies
hipError_t launch_bdpt(const WaveParams& P, const BdptArgsHost& A, uint32_t width, hipStream_t st);
// The wavefront form of the same chunk: the light and camera subpaths walked by extend and
// a subpath shade per bounce, the strategies' connections traced by connect.
// side[0..1]: two more streams and ev[0..2]: three events for the strategy kernels, which
// run side by side
hipError_t launch_bdpt_wave(const WaveParams& P, const BdptArgsHost& A, const KernelCfg& ce,
                            const KernelCfg& cc, uint32_t shade_grid, uint32_t part, hipStream_t st,
                            const hipStream_t side[2], const hipEvent_t ev[3]);
hipError_t launch_extend(const WaveParams& P, const KernelCfg& cfg, uint32_t bounce, bool batch,
                         hipStream_t st);
// max_blocks: the shade grid (workgroups of 256 that stride over the queue)
hipError_t launch_shade(const WaveParams& P, uint32_t capacity, uint32_t bounce, bool batch, hipStream_t st,
                        uint32_t max_blocks);
// The adjoint replay's shade (k_shade with gradient deposits), the first pass's per-path
// radiance (k_adj_total, per pixel) and the last bounce's pending deposits (k_adj_final).
hipError_t launch_shade_adjoint(const WaveParams& P, uint32_t capacity, uint32_t bounce, hipStream_t st,
                                uint32_t max_blocks);
hipError_t launch_adj_total(const WaveParams& P, hipStream_t st);
hipError_t launch_adj_final(const WaveParams& P, hipStream_t st);
hipError_t launch_connect(const WaveParams& P, const KernelCfg& cfg, uint32_t bounce, bool batch,
                          hipStream_t st);
hipError_t launch_finalize(const WaveParams& P, uint32_t capacity, bool batch, hipStream_t st);
hipError_t launch_advance(uint32_t* cur, hipStream_t st);
hipError_t launch_adjoint(const WaveParams& P, const AdjArgs& A, uint32_t width, hipStream_t st);
hipError_t launch_zero(uint32_t* p, uint32_t n, hipStream_t st);
hipError_t launch_lt_start(const WaveParams& P, hipStream_t st, uint32_t grid);
hipError_t launch_lt_shade(const WaveParams& P, uint32_t bounce, hipStream_t st, uint32_t grid);
hipError_t launch_lt_splat(const WaveParams& P, uint32_t stage, hipStream_t st, uint32_t grid);
hipError_t launch_lt_stats(const WaveParams& P, hipStream_t st);

// A context's lock, device and stream, for entry points outside api.hip.
std::mutex& ctx_mutex(dgpu_ctx* ctx);
int ctx_device(dgpu_ctx* ctx);
hipStream_t ctx_stream(dgpu_ctx* ctx);

}  // namespace dgpu
