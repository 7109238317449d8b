// api.hip -- the C-ABI of include/dortgpu.h: contexts, scene upload, intersection
// queries and the wavefront render loop.
//
// The render loop replaces PathRenderer::render (src/dort/path_renderer.cpp:11-21)
// and Renderer::iteration_tiled (src/dort/renderer.cpp:26-86): instead of tiles x
// iterations on a thread pool, the paths of [sample_begin, sample_end) x pixels are
// processed in chunks of `queue_tile` paths, each chunk running the bounce loop
// raygen -> (extend -> shade -> connect) x (max_depth + 1) -> finalize on one stream.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "bvh_build.hpp"
#include "dgpu_internal.hpp"
#include "status.hpp"

using namespace dgpu;


namespace {

#define HIP_TRY(x)                                                                   \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) return fail(DGPU_EDEVICE, "%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// Scratch of dgpu_trace_closest_noalloc, for callers that cannot allocate device memory:
// the persistent kernel's fetch counter (the first 256 bytes) and the traversal stack's
// spill slices.  One array per GPU, held from the library's first use until the process
// ends.  Host code never uses the array's name as an address: dgpu_create looks up each
// device's copy with hipGetSymbolAddress and keeps it in scratch_ptr.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;
constexpr size_t NOALLOC_COUNTER_BYTES = 256;
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex scratch_mu[NOALLOC_MAX_DEVICES];       // one call at a time per GPU
uint8_t* scratch_ptr[NOALLOC_MAX_DEVICES] = {};   // the array's address on each GPU

}  // namespace

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

struct dgpu_ctx {
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  // the render's shadow-ray tests run on a second stream of the same (default) priority,
  // each bounce's beside the next bounce's extension rays; created on first use
  hipStream_t stream2 = nullptr;
  hipStream_t stream3 = nullptr;  // the light tracer's odd stages' camera connections
  std::vector<hipEvent_t> ev_sync;  // ordering events (no timing), reused across renders
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::mutex mu;
  std::map<uint64_t, uint32_t> grid_cache;
  // render scratch, grown on demand
  std::map<std::string, DevBuf> bufs;
  // HIP events for per-launch profiling, reused across renders
  std::vector<hipEvent_t> ev_pool;

  hipEvent_t pool_event(size_t i) {
    while (ev_pool.size() <= i) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      ev_pool.push_back(e);
    }
    return ev_pool[i];
  }

  hipEvent_t sync_event(size_t i) {
    while (ev_sync.size() <= i) {
      hipEvent_t e = nullptr;
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
      ev_sync.push_back(e);
    }
    return ev_sync[i];
  }

  hipStream_t second_stream() {
    if (!stream2 && hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking) != hipSuccess) stream2 = nullptr;
    return stream2;
  }

  hipStream_t third_stream() {
    if (!stream3 && hipStreamCreateWithFlags(&stream3, hipStreamNonBlocking) != hipSuccess) stream3 = nullptr;
    return stream3;
  }

  ~dgpu_ctx() {
    (void)hipSetDevice(device);
    for (auto& kv : bufs) if (kv.second.p) (void)hipFree(kv.second.p);
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_sync) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream3) (void)hipStreamDestroy(stream3);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // Returns a device buffer of at least `bytes` under `name` (reused across calls).
  void* buf(const std::string& name, size_t bytes, dgpu_status* st) {
    DevBuf& b = bufs[name];
    if (b.bytes >= bytes && b.p) return b.p;
    if (b.p) { (void)hipFree(b.p); b.p = nullptr; b.bytes = 0; }
    size_t want = std::max<size_t>(bytes, 256);
    if (hipMalloc(&b.p, want) != hipSuccess) {
      b.p = nullptr;
      *st = fail(DGPU_ENOMEM, "hipMalloc(%zu) failed for %s", want, name.c_str());
      return nullptr;
    }
    b.bytes = want;
    return b.p;
  }

  // Paths per wavefront chunk when the scene sets none: as many as fit a quarter of
  // the free device memory, at most 16 GiB of queues and path state.  Bigger chunks
  // amortise the ramp-up and drain of every per-bounce launch (on the 2M-triangle
  // terrain the traversal kernels ran twice as fast at 64M paths as at 2M).
  // sets: shadow queues and pending records per path (2 when a bounce's NEE is added two
  // bounces later), each set with its own 16 GiB of the budget.
  uint64_t default_tile(uint32_t slots, uint32_t sets = 1) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    const uint64_t per_path = 120ull + 56ull * slots * sets;  // state, queues, hit, shadow queue
    uint64_t budget = std::min<uint64_t>((16ull << 30) * sets, free_b / 4);
    return std::max<uint64_t>(1ull << 20, budget / per_path);
  }

  // Paths per batch chunk: up to a quarter of the free device memory, at most 32 GiB
  // (148M paths; chunks larger than about 2^27 paths measured no faster).
  uint64_t batch_tile(uint32_t slots) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    const uint64_t per_path = 120ull + 56ull * slots;  // as default_tile's
    uint64_t budget = std::min<uint64_t>(32ull << 30, free_b / 4);
    return std::max<uint64_t>(1ull << 20, budget / per_path);
  }

  // Workgroups of the strided shade kernel: enough to fill every CU.
  uint32_t shade_blocks(bool batch, bool all, bool rough) {
    int& cached = shade_grid_cache[batch][all][rough];
    if (cached == 0) {
      int nb = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, shade_kernel_ptr(batch, all, rough), SHADE_BLOCK, 0) !=
              hipSuccess ||
          nb <= 0)
        nb = 2;
      cached = nb * num_cus * 4;
    }
    return (uint32_t)cached;
  }
  int shade_grid_cache[2][2][2] = {};  // [batch][sample_all][rough]

  // Workgroups of the adjoint replay's strided shade kernel.
  uint32_t shade_blocks_adjoint(bool all, bool rough) {
    int& cached = shade_grid_adj_cache[all][rough];
    if (cached == 0) {
      int nb = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, shade_kernel_ptr(false, all, rough, true), SHADE_BLOCK,
                                                       0) != hipSuccess ||
          nb <= 0)
        nb = 2;
      cached = nb * num_cus * 4;
    }
    return (uint32_t)cached;
  }
  int shade_grid_adj_cache[2][2] = {};  // [sample_all][rough]

  uint32_t grid_for(int which, const KernelCfg& c) {
    uint64_t key = ((uint64_t)which << 48) | ((uint64_t)c.width << 32) | ((uint64_t)c.lds << 16) | c.block;
    auto it = grid_cache.find(key);
    if (it != grid_cache.end()) return it->second;
    int nb = 0;
    const void* k = trace_kernel_ptr(which, c.width, c.lds, c.block);
    if (!k || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, (int)c.block, 0) != hipSuccess || nb <= 0)
      nb = 4;
    uint32_t g = (uint32_t)nb * (uint32_t)num_cus;
    grid_cache[key] = g;
    return g;
  }
};

std::mutex& dgpu::ctx_mutex(dgpu_ctx* ctx) { return ctx->mu; }
int dgpu::ctx_device(dgpu_ctx* ctx) { return ctx->device; }
hipStream_t dgpu::ctx_stream(dgpu_ctx* ctx) { return ctx->stream; }


struct dgpu_scene {
  dgpu_ctx* ctx = nullptr;
  DevScene S;
  std::vector<void*> allocs;
  KernelCfg cfg;
  uint32_t queue_tile = 0;
  uint32_t closest_grid = 0;  // persistent grid of the closest-hit intersection kernel
  uint32_t stack_cap = MAX_STACK;
  uint32_t num_area = 0;
  dgpu_scene_info info;
  std::vector<uint4> tex_info_host;   // (w, h, texel offset, mapping)
  std::vector<DevMaterial> mats_host;

  ~dgpu_scene() {
    if (ctx) (void)hipSetDevice(ctx->device);
    for (void* p : allocs) (void)hipFree(p);
  }

  template <class T>
  dgpu_status upload(const T* host, size_t count, const T** out) {
    *out = nullptr;
    if (count == 0) return DGPU_OK;
    void* d = nullptr;
    if (hipMalloc(&d, count * sizeof(T)) != hipSuccess) return fail(DGPU_ENOMEM, "hipMalloc(%zu) failed", count * sizeof(T));
    allocs.push_back(d);
    HIP_TRY(hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice));
    *out = (const T*)d;
    return DGPU_OK;
  }
};

namespace {

dgpu_status validate_desc(const dgpu_scene_desc* d) {
  if (!d) return fail(DGPU_EINVAL, "null scene descriptor");
  if (d->num_triangles == 0) return fail(DGPU_EINVAL, "scene has no triangles");
  if (!d->positions || !d->indices || !d->tri_material)
    return fail(DGPU_EINVAL, "positions, indices and tri_material are required");
  if (d->num_materials == 0 || !d->materials) return fail(DGPU_EINVAL, "scene has no materials");
  for (size_t i = 0; i < 3 * (size_t)d->num_triangles; ++i)
    if (d->indices[i] >= d->num_vertices) return fail(DGPU_EINVAL, "index %zu out of range", i);
  for (uint32_t t = 0; t < d->num_triangles; ++t)
    if (d->tri_material[t] >= d->num_materials) return fail(DGPU_EINVAL, "triangle %u: bad material", t);
  for (size_t i = 0; i < 3 * (size_t)d->num_vertices; ++i)
    if (!std::isfinite(d->positions[i])) return fail(DGPU_EINVAL, "non-finite vertex coordinate");
  for (uint32_t m = 0; m < d->num_materials; ++m) {
    const dgpu_material& mt = d->materials[m];
    if (mt.type > DGPU_MAT_ROUGH_DIELECTRIC)
      return fail(DGPU_EUNSUPPORTED, "material %u: unsupported type %u", m, mt.type);
    if (mt.type == DGPU_MAT_ROUGH_DIELECTRIC && !(mt.roughness >= 0.f && std::isfinite(mt.roughness)))
      return fail(DGPU_EINVAL, "material %u: rough dielectric roughness must be finite and >= 0", m);
    if (mt.type == DGPU_MAT_LAMBERT && mt.albedo_tex >= (int32_t)d->num_textures)
      return fail(DGPU_EINVAL, "material %u: texture %d out of range", m, mt.albedo_tex);
    if (mt.type == DGPU_MAT_GGX_CONDUCTOR && !(mt.roughness > 0.f))
      return fail(DGPU_EINVAL, "material %u: GGX roughness must be > 0", m);
  }
  for (uint32_t l = 0; l < d->num_area_lights; ++l)
    if (d->area_lights[l].triangle >= d->num_triangles) return fail(DGPU_EINVAL, "area light %u: bad triangle", l);
  if (d->env && (!d->env->rgb || d->env->width == 0 || d->env->height == 0))
    return fail(DGPU_EINVAL, "environment map without pixels");
  for (uint32_t t = 0; t < d->num_textures; ++t) {
    const dgpu_texture& tx = d->textures[t];
    if ((!tx.rgb && !tx.rgb8) || tx.width == 0 || tx.height == 0)
      return fail(DGPU_EINVAL, "texture %u without texels", t);
    if (tx.mapping > DGPU_TEXMAP_XY) return fail(DGPU_EUNSUPPORTED, "texture %u: unknown mapping %u", t, tx.mapping);
    if (tx.mapping == DGPU_TEXMAP_XY)
      for (int k = 0; k < 8; ++k)
        if (!std::isfinite(tx.world_to_texture[k])) return fail(DGPU_EINVAL, "texture %u: non-finite transform", t);
  }
  if (!(d->camera.fov > 0.f && d->camera.fov < 3.14159f)) return fail(DGPU_EINVAL, "bad camera fov");
  return DGPU_OK;
}

DevMaterial to_dev_material(const dgpu_material& a) {
  DevMaterial b;
  memset(&b, 0, sizeof b);
  b.type = a.type;
  b.tex = a.type == DGPU_MAT_LAMBERT ? a.albedo_tex : -1;
  b.roughness = a.roughness;
  b.ior_in = a.ior_inside;
  b.albedo = make_float4(a.albedo[0], a.albedo[1], a.albedo[2], a.ior_outside);
  b.eta = make_float4(a.eta[0], a.eta[1], a.eta[2], 0.f);
  b.k = make_float4(a.k[0], a.k[1], a.k[2], 0.f);
  b.refl = make_float4(a.reflect_tint[0], a.reflect_tint[1], a.reflect_tint[2], 0.f);
  b.trans = make_float4(a.transmit_tint[0], a.transmit_tint[1], a.transmit_tint[2], 0.f);
  return b;
}

}  // namespace

extern "C" {

uint32_t dgpu_abi_version(void) { return DGPU_ABI_VERSION; }

dgpu_status dgpu_create(int device, dgpu_ctx** out) {
  if (!out) return fail(DGPU_EINVAL, "null output");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(DGPU_EDEVICE, "no HIP device");
  if (device < 0 || device >= n) return fail(DGPU_EINVAL, "device %d out of range (%d devices)", device, n);
  dgpu_ctx* c = new dgpu_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete c; return fail(DGPU_EDEVICE, "hipSetDevice failed"); }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return fail(DGPU_EDEVICE, "hipGetDeviceProperties failed"); }
  c->num_cus = prop.multiProcessorCount;
  if (device < NOALLOC_MAX_DEVICES) {
    std::lock_guard<std::mutex> lk(scratch_mu[device]);
    if (!scratch_ptr[device]) {
      void* p = nullptr;
      hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        delete c;
        return fail(DGPU_EDEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      }
      scratch_ptr[device] = (uint8_t*)p;
    }
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return fail(DGPU_EDEVICE, "stream/event creation failed");
  }
  *out = c;
  return DGPU_OK;
}

void dgpu_destroy(dgpu_ctx* ctx) { delete ctx; }

dgpu_status dgpu_device_alloc(dgpu_ctx* ctx, size_t bytes, void** out) {
  if (!ctx || !out) return fail(DGPU_EINVAL, "null context or output");
  *out = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) {
    *out = nullptr;
    return fail(DGPU_ENOMEM, "hipMalloc(%zu) failed", bytes);
  }
  return DGPU_OK;
}

dgpu_status dgpu_device_free(dgpu_ctx* ctx, void* ptr) {
  if (!ctx) return fail(DGPU_EINVAL, "null context");
  HIP_TRY(hipSetDevice(ctx->device));
  if (ptr) HIP_TRY(hipFree(ptr));
  return DGPU_OK;
}

dgpu_status dgpu_copy(dgpu_ctx* ctx, void* dst, const void* src, size_t bytes, int kind) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(DGPU_EINVAL, "null argument");
  if (kind < 0 || kind > 2) return fail(DGPU_EINVAL, "bad copy kind %d", kind);
  if (bytes == 0) return DGPU_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return DGPU_OK;
}

dgpu_status dgpu_scene_create(dgpu_ctx* ctx, const dgpu_scene_desc* d, const dgpu_bvh_opts* opts,
                              dgpu_scene** out) {
  if (!ctx || !out) return fail(DGPU_EINVAL, "null context or output");
  *out = nullptr;
  dgpu_status st = validate_desc(d);
  if (st != DGPU_OK) return st;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  dgpu_bvh_opts o{};
  if (opts) o = *opts;
  KernelCfg cfg;
  cfg.width = o.width ? o.width : 4;
  cfg.lds = o.stack_lds ? o.stack_lds : 16;
  cfg.block = o.block_size ? o.block_size : 128;
  cfg.grid = 0;
  if (!variant_supported(cfg.width, cfg.lds, cfg.block))
    return fail(DGPU_EUNSUPPORTED, "no kernel variant for width %u, stack %u, block %u", cfg.width, cfg.lds, cfg.block);

  auto t0 = std::chrono::steady_clock::now();
  dgpu_host::BvhParams bp;
  bp.width = cfg.width;
  bp.max_leaf_size = o.max_leaf_size ? o.max_leaf_size : 4;
  bp.sah_bins = o.sah_bins ? o.sah_bins : 16;
  bp.node_cost = o.sah_node_cost > 0.f ? o.sah_node_cost : 1.f;
  {
    // Box padding for the path tracer's FMA slab test: M bounds every coordinate of the
    // scene and of the camera position (where all of its rays start); the test's error
    // is below 3 * 2^-24 * M, the padding is M * 2^-21.
    float m = 0.f;
    for (size_t t = 0; t < 3 * (size_t)d->num_triangles; ++t) {
      const float* p = d->positions + 3 * (size_t)d->indices[t];
      m = std::max(m, std::max(std::fabs(p[0]), std::max(std::fabs(p[1]), std::fabs(p[2]))));
    }
    const float* c2w = d->camera.camera_to_world;
    m = std::max(m, std::max(std::fabs(c2w[3]), std::max(std::fabs(c2w[7]), std::fabs(c2w[11]))));
    bp.box_pad = std::isfinite(m) ? std::ldexp(m, -21) : 0.f;
  }
  if (!(o.sah_node_cost >= 0.f) || o.sah_node_cost > 64.f) return fail(DGPU_EINVAL, "sah_node_cost must be in [0, 64]");
  if (bp.max_leaf_size > 16) return fail(DGPU_EINVAL, "max_leaf_size must be <= 16");
  dgpu_host::BuiltBvh bvh;
  try {
    bvh = dgpu_host::build_bvh(d->positions, d->num_vertices, d->indices, d->num_triangles, bp);
  } catch (const std::exception& e) {
    return fail(DGPU_EINVAL, "BVH build failed: %s", e.what());
  }
  auto t1 = std::chrono::steady_clock::now();

  dgpu_scene* s = new dgpu_scene();
  s->ctx = ctx;
  s->cfg = cfg;
  s->closest_grid = ctx->grid_for(0, cfg);  // kept for dgpu_trace_closest_noalloc, which allocates nothing
  s->queue_tile = o.queue_tile;
  DevScene& S = s->S;
  memset(&S, 0, sizeof S);
  S.width = bvh.width;
  S.num_tris = d->num_triangles;
  s->stack_cap = (bvh.width - 1) * (bvh.depth + 1) + 2;
  if (s->stack_cap < cfg.lds) s->stack_cap = cfg.lds;
  S.stack_cap = s->stack_cap;
#define UP(host, count, dst)                                                  \
  do {                                                                        \
    dgpu_status us_ = s->upload(host, count, &dst);                           \
    if (us_ != DGPU_OK) { delete s; return us_; }                             \
  } while (0)
  UP((const float4*)bvh.nodes.data(), bvh.nodes.size() / 4, S.nodes);

  // scene bounds (Scene::bounds -> centroid / radius, lua_builder.cpp:124-128)
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t t = 0; t < 3 * (size_t)d->num_triangles; ++t) {
    const float* p = d->positions + 3 * (size_t)d->indices[t];
    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
  }
  float radius = 0.5f * std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                  (hi[2] - lo[2]) * (hi[2] - lo[2]));

  // per-triangle uvs, in leaf order
  if (d->uvs) {
    std::vector<float2> tuv(3 * (size_t)d->num_triangles);
    for (size_t k = 0; k < (size_t)d->num_triangles; ++k) {
      uint32_t t;
      memcpy(&t, &bvh.tris[12 * k + 3], 4);
      for (int j = 0; j < 3; ++j) {
        uint32_t v = d->indices[3 * (size_t)t + j];
        tuv[3 * k + j] = make_float2(d->uvs[2 * (size_t)v], d->uvs[2 * (size_t)v + 1]);
      }
    }
    UP(tuv.data(), tuv.size(), S.tri_uv);
  }
  // materials
  std::vector<DevMaterial> mats(d->num_materials);
  for (uint32_t m = 0; m < d->num_materials; ++m) mats[m] = to_dev_material(d->materials[m]);
  UP(mats.data(), mats.size(), S.mats);
  S.num_mats = d->num_materials;
  for (const DevMaterial& m : mats) S.has_rough |= m.type == DGPU_MAT_ROUGH_DIELECTRIC ? 1u : 0u;
  s->mats_host = mats;

  // lights, in dort's add order: area lights, environment, infinite
  std::vector<DevLight> lights;
  std::vector<int32_t> tri_light(d->num_triangles, -1);
  std::vector<float> powers;
  for (uint32_t l = 0; l < d->num_area_lights; ++l) {
    const dgpu_area_light& al = d->area_lights[l];
    const float* p0 = d->positions + 3 * (size_t)d->indices[3 * (size_t)al.triangle];
    const float* p1 = d->positions + 3 * (size_t)d->indices[3 * (size_t)al.triangle + 1];
    const float* p2 = d->positions + 3 * (size_t)d->indices[3 * (size_t)al.triangle + 2];
    float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    float cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
    float area = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
    DevLight L;
    memset(&L, 0, sizeof L);
    L.kind = LK_AREA;
    L.tri = bvh.tri_leaf_of[al.triangle];
    L.area = area;
    L.radiance = make_float4(al.radiance[0], al.radiance[1], al.radiance[2], 0.f);
    tri_light[al.triangle] = (int32_t)lights.size();
    lights.push_back(L);
    // DiffuseLight::approximate_power (diffuse_light.cpp:117-119)
    powers.push_back((area * al.radiance[0] + area * al.radiance[1] + area * al.radiance[2]) / 3.f);
  }
  S.env_light = -1;
  S.inf_light = -1;
  if (d->env) {
    const dgpu_env_light& e = *d->env;
    uint32_t w = e.width, h = e.height;
    S.env_w = w;
    S.env_h = h;
    // frame (environment_light.cpp:6-13)
    V3 up = v3(e.up[0], e.up[1], e.up[2]), fwd = v3(e.fwd[0], e.fwd[1], e.fwd[2]);
    S.env_up = normalize(up);
    S.env_t = normalize(cross(up, fwd));
    S.env_s = cross(S.env_up, S.env_t);
    std::vector<float4> rgb((size_t)w * h);
    std::vector<float> values((size_t)w * h);
    double vs[3] = {0, 0, 0};
    float vsum[3] = {0.f, 0.f, 0.f}, wsum = 0.f;
    (void)vs;
    for (uint32_t y = 0; y < h; ++y) {
      float theta = (float(y) + 0.5f) / float(h) * PI_F;
      float ct = clampf(std::cos(theta), 0.f, 1.f);
      for (uint32_t x = 0; x < w; ++x) {
        const float* px = e.rgb + 3 * ((size_t)y * w + x);
        rgb[(size_t)y * w + x] = make_float4(px[0] * e.scale[0], px[1] * e.scale[1], px[2] * e.scale[2], 0.f);  // .w below
        values[(size_t)y * w + x] = std::fabs((px[0] + px[1] + px[2]) / 3.f * ct);
        for (int k = 0; k < 3; ++k) vsum[k] += px[k] * ct;
        wsum += ct;
      }
    }
    // PiecewiseDistrib2d (piecewise_distrib_2d.cpp:5-34)
    std::vector<float> pdfs((size_t)w * h), marg(h + 1), cond((size_t)(w + 1) * h);
    float sum = 0.f;
    marg[0] = 0.f;
    for (uint32_t y = 0; y < h; ++y) {
      float ys = 0.f;
      cond[(size_t)(w + 1) * y] = 0.f;
      for (uint32_t x = 0; x < w; ++x) {
        ys += values[(size_t)y * w + x];
        cond[(size_t)(w + 1) * y + x + 1] = ys;
      }
      float inv = 1.f / ys;
      for (uint32_t x = 0; x <= w; ++x) cond[(size_t)(w + 1) * y + x] *= inv;
      sum += ys;
      marg[y + 1] = sum;
    }
    float inv_sum = 1.f / sum;
    for (uint32_t y = 0; y <= h; ++y) marg[y] *= inv_sum;
    for (size_t i = 0; i < (size_t)w * h; ++i) pdfs[i] = values[i] * inv_sum;
    for (size_t i = 0; i < (size_t)w * h; ++i) rgb[i].w = pdfs[i];  // radiance and probability in one fetch
    if (!(sum > 0.f)) { delete s; return fail(DGPU_EINVAL, "environment map has no energy in the upper hemisphere"); }
    UP(rgb.data(), rgb.size(), S.env_rgb);
    UP(marg.data(), marg.size(), S.env_marg);
    UP(cond.data(), cond.size(), S.env_cond);
    // guide tables: for u = g / G, the bin the bisection would return (a start hint;
    // the device scan corrects it in either direction)
    auto guide = [](const float* cdf, uint32_t size, uint32_t G, uint32_t* out) {
      uint32_t k = 0;  // entries of cdf[0 .. size-2] below u
      for (uint32_t g = 0; g < G; ++g) {
        float u = float(g) / float(G);
        while (k < size - 1 && cdf[k] < u) ++k;
        out[g] = k == 0 ? 0u : std::min(k - 1u, size - 2u);
      }
    };
    std::vector<uint32_t> mg(h), cg((size_t)w * h);
    guide(marg.data(), h + 1, h, mg.data());
    for (uint32_t y = 0; y < h; ++y) guide(cond.data() + (size_t)(w + 1) * y, w + 1, w, cg.data() + (size_t)w * y);
    UP(mg.data(), mg.size(), S.env_marg_guide);
    UP(cg.data(), cg.size(), S.env_cond_guide);
    DevLight L;
    memset(&L, 0, sizeof L);
    L.kind = LK_ENV;
    S.env_light = (int32_t)lights.size();
    lights.push_back(L);
    // EnvironmentLight::approximate_power (environment_light.cpp:130-132)
    float avg[3];
    for (int k = 0; k < 3; ++k) avg[k] = e.scale[k] * vsum[k] / wsum;
    float a = PI_F * radius * radius;
    powers.push_back((a * avg[0] + a * avg[1] + a * avg[2]) / 3.f);
  }
  if (d->infinite_radiance) {
    DevLight L;
    memset(&L, 0, sizeof L);
    L.kind = LK_INF;
    L.radiance = make_float4(d->infinite_radiance[0], d->infinite_radiance[1], d->infinite_radiance[2], 0.f);
    S.has_inf = 1;
    S.inf_rad = v3(d->infinite_radiance[0], d->infinite_radiance[1], d->infinite_radiance[2]);
    S.inf_light = (int32_t)lights.size();
    lights.push_back(L);
    float a = PI_F * radius * radius;  // InfiniteLight::approximate_power
    powers.push_back((a * L.radiance.x + a * L.radiance.y + a * L.radiance.z) / 3.f);
  }
  // the background lights' own DiscreteDistrib1d (environment first, then infinite;
  // discrete_distrib_1d.cpp:4-20)
  {
    float pe = S.env_light >= 0 ? powers[S.env_light] : 0.f;
    float pi = S.inf_light >= 0 ? powers[S.inf_light] : 0.f;
    float sum = pe + pi, inv = sum == 0.f ? 0.f : 1.f / sum;
    S.bg_env_prob = S.env_light >= 0 ? pe * inv : 0.f;
    S.bg_inf_prob = S.inf_light >= 0 ? (pe + pi) * inv - (S.env_light >= 0 ? pe * inv : 0.f) : 0.f;
  }
  // DiscreteDistrib1d over light powers (scene.cpp:20-28, discrete_distrib_1d.cpp:4-20)
  std::vector<float> cdf;
  float acc = 0.f;
  for (float p : powers) { cdf.push_back(acc); acc += p; }
  cdf.push_back(acc);
  float inv = acc == 0.f ? 0.f : 1.f / acc;
  for (float& c : cdf) c *= inv;
  S.num_lights = (uint32_t)lights.size();
  s->num_area = d->num_area_lights;
  if (!lights.empty()) {
    UP(lights.data(), lights.size(), S.lights);
    UP(cdf.data(), cdf.size(), S.light_cdf);
  }
  // leaf-order triangle records: v0.w scene id (from the builder), v1.w material,
  // v2.w light index or -1, so a hit reaches its attributes in one fetch
  for (size_t k = 0; k < (size_t)d->num_triangles; ++k) {
    uint32_t t;
    memcpy(&t, &bvh.tris[12 * k + 3], 4);
    uint32_t m = d->tri_material[t];
    uint32_t l = (uint32_t)tri_light[t];
    memcpy(&bvh.tris[12 * k + 7], &m, 4);
    memcpy(&bvh.tris[12 * k + 11], &l, 4);
  }
  UP((const float4*)bvh.tris.data(), bvh.tris.size() / 4, S.tris);
  // textures
  if (d->num_textures) {
    std::vector<float4> texels, xform;
    std::vector<uint4> info;
    for (uint32_t t = 0; t < d->num_textures; ++t) {
      const dgpu_texture& tx = d->textures[t];
      info.push_back(make_uint4(tx.width, tx.height, (uint32_t)texels.size(), tx.mapping));
      for (size_t i = 0; i < (size_t)tx.width * tx.height; ++i) {
        if (tx.rgb) {
          texels.push_back(make_float4(tx.rgb[3 * i], tx.rgb[3 * i + 1], tx.rgb[3 * i + 2], 0.f));
        } else {  // PixelRgb8::to_rgb: k * 2^-8 (image.hpp:66-71), exact in fp32
          const uint8_t* c = tx.rgb8 + 3 * i;
          texels.push_back(make_float4(std::ldexp((float)c[0], -8), std::ldexp((float)c[1], -8),
                                       std::ldexp((float)c[2], -8), 0.f));
        }
      }
      // rows 0 and 1 of world_to_texture: the texture's (x, y) of a world point
      const float* m = tx.world_to_texture;
      bool xy = tx.mapping == DGPU_TEXMAP_XY;
      xform.push_back(xy ? make_float4(m[0], m[1], m[2], m[3]) : make_float4(0.f, 0.f, 0.f, 0.f));
      xform.push_back(xy ? make_float4(m[4], m[5], m[6], m[7]) : make_float4(0.f, 0.f, 0.f, 0.f));
    }
    UP(texels.data(), texels.size(), S.tex_data);
    UP(info.data(), info.size(), S.tex_info);
    UP(xform.data(), xform.size(), S.tex_xform);
    S.num_tex = d->num_textures;
    s->tex_info_host = info;
  }
#undef UP
  for (int i = 0; i < 12; ++i) S.c2w[i] = d->camera.camera_to_world[i];
  {
    // world-to-camera: the inverse of the affine camera_to_world (in double), as dort's
    // Transform keeps the inverse matrix for apply_inv
    const float* m = d->camera.camera_to_world;
    double a[3][3] = {{m[0], m[1], m[2]}, {m[4], m[5], m[6]}, {m[8], m[9], m[10]}};
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                 a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    if (!(det != 0.0) || !std::isfinite(det)) { delete s; return fail(DGPU_EINVAL, "camera transform is singular"); }
    double inv[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
        inv[r][c] = (a[r1][c1] * a[r2][c2] - a[r1][c2] * a[r2][c1]) / det;
      }
    double t[3] = {m[3], m[7], m[11]};
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) S.w2c[4 * r + c] = (float)inv[r][c];
      S.w2c[4 * r + 3] = (float)-(inv[r][0] * t[0] + inv[r][1] * t[1] + inv[r][2] * t[2]);
    }
  }
  S.centroid = v3(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]));
  S.radius = radius;
  S.project_dim = 2.f * std::tan(0.5f * d->camera.fov);
  auto t2 = std::chrono::steady_clock::now();

  s->info.nodes = bvh.num_nodes;
  s->info.leaves = bvh.num_leaves;
  s->info.depth = bvh.depth;
  s->info.width = bvh.width;
  s->info.node_bytes = (uint64_t)bvh.nodes.size() * 4;
  s->info.triangle_bytes = (uint64_t)bvh.tris.size() * 4;
  s->info.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  s->info.upload_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
  *out = s;
  return DGPU_OK;
}

void dgpu_scene_destroy(dgpu_scene* s) { delete s; }

dgpu_status dgpu_scene_get_info(const dgpu_scene* s, dgpu_scene_info* out) {
  if (!s || !out) return fail(DGPU_EINVAL, "null scene or output");
  *out = s->info;
  return DGPU_OK;
}

}  // extern "C"

namespace {

dgpu_status trace_common(dgpu_scene* s, const dgpu_rays* rays, uint32_t n, dgpu_hits* hits,
                         uint32_t* occluded, uint32_t flags, double* kernel_ms) {
  if (!s || !rays) return fail(DGPU_EINVAL, "null scene or rays");
  if (!occluded && (!hits || !hits->t || !hits->prim || !hits->b1 || !hits->b2))
    return fail(DGPU_EINVAL, "null hit buffers");
  const void* rp[8] = {rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
  for (const void* p : rp) if (!p) return fail(DGPU_EINVAL, "null ray component");
  if (kernel_ms) *kernel_ms = 0.0;
  if (n == 0) return DGPU_OK;
  dgpu_ctx* ctx = s->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  dgpu_status st = DGPU_OK;
  KernelCfg cfg = s->cfg;
  cfg.grid = ctx->grid_for(occluded ? 1 : 0, cfg);
  uint32_t* fetch = (uint32_t*)ctx->buf("fetch", 64, &st);
  uint32_t* spill = (uint32_t*)ctx->buf("spill", (size_t)cfg.grid * cfg.block * s->stack_cap * 4, &st);
  if (st != DGPU_OK) return st;
  dgpu_rays dr = *rays;
  dgpu_hits dh{nullptr, nullptr, nullptr, nullptr};
  uint32_t* docc = occluded;
  bool dev = flags & DGPU_BUF_DEVICE;
  hipStream_t stream = ctx->stream;
  if (!dev) {
    float* in = (float*)ctx->buf("trace_in", (size_t)n * 8 * 4, &st);
    float* outb = (float*)ctx->buf("trace_out", (size_t)n * 4 * 4, &st);
    if (st != DGPU_OK) return st;
    const float* comps[8] = {rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
    const float** dst[8] = {&dr.ox, &dr.oy, &dr.oz, &dr.dx, &dr.dy, &dr.dz, &dr.tmin, &dr.tmax};
    for (int k = 0; k < 8; ++k) {
      HIP_TRY(hipMemcpyAsync(in + (size_t)k * n, comps[k], (size_t)n * 4, hipMemcpyHostToDevice, stream));
      *dst[k] = in + (size_t)k * n;
    }
    if (occluded) docc = (uint32_t*)outb;
    else dh = dgpu_hits{outb, (int32_t*)(outb + n), outb + 2 * (size_t)n, outb + 3 * (size_t)n};
  } else if (hits) {
    dh = *hits;
  }
  HIP_TRY(hipEventRecord(ctx->ev0, stream));
  HIP_TRY(launch_trace_api(s->S, cfg, dr, occluded ? nullptr : &dh, docc, n, fetch, spill,
                           (flags & DGPU_TRACE_BOUNDED) != 0, stream));
  HIP_TRY(hipEventRecord(ctx->ev1, stream));
  if (!dev) {
    if (occluded) {
      HIP_TRY(hipMemcpyAsync(occluded, docc, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
    } else {
      HIP_TRY(hipMemcpyAsync(hits->t, dh.t, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipMemcpyAsync(hits->prim, dh.prim, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipMemcpyAsync(hits->b1, dh.b1, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipMemcpyAsync(hits->b2, dh.b2, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
    }
  }
  HIP_TRY(hipStreamSynchronize(stream));
  if (kernel_ms) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *kernel_ms = ms;
  }
  return DGPU_OK;
}

}  // namespace

extern "C" {

dgpu_status dgpu_trace_closest(dgpu_scene* s, const dgpu_rays* rays, uint32_t n, dgpu_hits* out,
                               uint32_t flags, double* kernel_ms) {
  if (!out) return fail(DGPU_EINVAL, "null hits");
  return trace_common(s, rays, n, out, nullptr, flags, kernel_ms);
}

dgpu_status dgpu_trace_closest_noalloc(dgpu_scene* s, const dgpu_rays* rays, uint32_t n, dgpu_hits* out,
                                       uint32_t flags, double* kernel_ms) {
  if (!s || !rays || !out) return fail(DGPU_EINVAL, "null scene, rays or hits");
  if (!out->t || !out->prim || !out->b1 || !out->b2) return fail(DGPU_EINVAL, "null hit buffers");
  const void* rp[8] = {rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
  for (const void* p : rp) if (!p) return fail(DGPU_EINVAL, "null ray component");
  if (!(flags & DGPU_BUF_DEVICE))
    return fail(DGPU_EINVAL, "dgpu_trace_closest_noalloc takes device buffers only (DGPU_BUF_DEVICE)");
  if (flags & ~(uint32_t)(DGPU_BUF_DEVICE | DGPU_TRACE_BOUNDED)) return fail(DGPU_EINVAL, "unknown flags 0x%x", flags);
  if (kernel_ms) *kernel_ms = 0.0;
  if (n == 0) return DGPU_OK;
  dgpu_ctx* ctx = s->ctx;
  if (ctx->device >= NOALLOC_MAX_DEVICES || !scratch_ptr[ctx->device])
    return fail(DGPU_EUNSUPPORTED, "no scratch array on device %d", ctx->device);
  // The spill slices of G threads take G * (stack entries beyond LDS) words after the
  // counter area; the persistent grid shrinks to what the array holds.
  KernelCfg cfg = s->cfg;
  const size_t spill_words = s->stack_cap > cfg.lds ? s->stack_cap - cfg.lds : 0;
  const size_t block_bytes = (size_t)cfg.block * spill_words * 4;
  const size_t avail = NOALLOC_SCRATCH_BYTES - NOALLOC_COUNTER_BYTES;
  const size_t max_blocks = block_bytes ? avail / block_bytes : (size_t)s->closest_grid;
  if (max_blocks == 0)
    return fail(DGPU_EUNSUPPORTED, "traversal stack of %u entries needs %zu bytes per workgroup, above the %zu-byte "
                "scratch", s->stack_cap, block_bytes, avail);
  cfg.grid = (uint32_t)std::min<size_t>(s->closest_grid, max_blocks);
  std::lock_guard<std::mutex> lk(ctx->mu);
  std::lock_guard<std::mutex> dlk(scratch_mu[ctx->device]);
  HIP_TRY(hipSetDevice(ctx->device));
  uint32_t* fetch = (uint32_t*)scratch_ptr[ctx->device];
  uint32_t* spill = (uint32_t*)(scratch_ptr[ctx->device] + NOALLOC_COUNTER_BYTES);
  hipStream_t stream = ctx->stream;
  HIP_TRY(hipEventRecord(ctx->ev0, stream));
  HIP_TRY(launch_trace_api(s->S, cfg, *rays, out, nullptr, n, fetch, spill, (flags & DGPU_TRACE_BOUNDED) != 0,
                           stream));
  HIP_TRY(hipEventRecord(ctx->ev1, stream));
  HIP_TRY(hipStreamSynchronize(stream));  // the device's scratch is free again once this returns
  if (kernel_ms) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *kernel_ms = ms;
  }
  return DGPU_OK;
}

dgpu_status dgpu_trace_any(dgpu_scene* s, const dgpu_rays* rays, uint32_t n, uint32_t* occluded,
                           uint32_t flags, double* kernel_ms) {
  if (!occluded) return fail(DGPU_EINVAL, "null occluded buffer");
  return trace_common(s, rays, n, nullptr, occluded, flags, kernel_ms);
}

dgpu_status dgpu_render(dgpu_scene* s, const dgpu_render_opts* o, float* film_rgb_sum,
                        float* film_weight, dgpu_stats* out_stats, const volatile int* cancel) {
  if (!s || !o) return fail(DGPU_EINVAL, "null scene or options");
  if (!film_rgb_sum || !film_weight) return fail(DGPU_EINVAL, "null film buffers");
  if (o->x_res == 0 || o->y_res == 0) return fail(DGPU_EINVAL, "empty film");
  if (o->sample_end < o->sample_begin) return fail(DGPU_EINVAL, "sample_end < sample_begin");
  if (o->max_depth + 2 >= MAX_BOUNCES) return fail(DGPU_EINVAL, "max_depth must be < %u", MAX_BOUNCES - 2);
  if (o->direct_strategy > DGPU_DIRECT_LIGHT) return fail(DGPU_EINVAL, "bad direct_strategy");
  if ((uint64_t)o->x_res * o->y_res > 0xFFFFFFFFull || o->x_res >= (1u << 28)) return fail(DGPU_EINVAL, "film too large");
  if (o->sample_all_lights && s->S.num_lights > MAX_SLOT_LIGHTS)
    return fail(DGPU_EUNSUPPORTED, "sample_all_lights supports at most %u lights", MAX_SLOT_LIGHTS);
  if (out_stats) memset(out_stats, 0, sizeof *out_stats);
  dgpu_ctx* ctx = s->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t stream = ctx->stream;
  dgpu_status st = DGPU_OK;
  const uint32_t npix = o->x_res * o->y_res;
  const uint64_t total = (uint64_t)npix * (o->sample_end - o->sample_begin);
  uint32_t slots = o->sample_all_lights ? 2 * std::max(1u, s->S.num_lights) : 2;
  uint64_t tile = s->queue_tile ? s->queue_tile : ctx->default_tile(slots, 2);
  // q0 + n_chunk must stay below 2^32 (the kernels' 32-bit path positions)
  uint32_t C = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(total ? total : 1, tile),
                                            std::min<uint64_t>(0xFFFFFFFFull / slots, std::max<uint64_t>(1, 0xFFFFFFFFull - npix)));
  const uint32_t shade_grid = ctx->shade_blocks(false, o->sample_all_lights != 0, s->S.has_rough != 0);

  KernelCfg ce = s->cfg, cc = s->cfg;
  ce.grid = ctx->grid_for(2, ce);
  cc.grid = ctx->grid_for(3, cc);
  uint32_t maxgrid = std::max(ce.grid, cc.grid);
  WaveParams P;
  memset(&P, 0, sizeof P);
  P.S = s->S;
  P.xr = o->x_res; P.yr = o->y_res; P.npix = npix;
  P.ord = make_path_order(o->x_res, o->y_res);
  P.min_depth = o->min_depth; P.max_depth = o->max_depth; P.rr_depth = o->rr_depth;
  P.only_direct = o->only_direct; P.sample_all = o->sample_all_lights; P.strategy = o->direct_strategy;
  P.slots = slots;
  P.rough = s->S.has_rough;
  dgpu_rng_key(o->seed, o->scene_id, &P.k0, &P.k1);
  P.sample_begin = o->sample_begin;
  P.thr = (float4*)ctx->buf("thr", (size_t)C * 16, &st);
  P.rad = (float4*)ctx->buf("rad", (size_t)C * 16, &st);
  P.pend = (float4*)ctx->buf("pend", (size_t)C * slots * 16, &st);
  P.qo[0] = (float4*)ctx->buf("qo0", (size_t)C * 16, &st);
  P.qo[1] = (float4*)ctx->buf("qo1", (size_t)C * 16, &st);
  P.qd[0] = (float4*)ctx->buf("qd0", (size_t)C * 16, &st);
  P.qd[1] = (float4*)ctx->buf("qd1", (size_t)C * 16, &st);
  P.qpath[0] = (uint32_t*)ctx->buf("qp0", (size_t)C * 4, &st);
  P.qpath[1] = (uint32_t*)ctx->buf("qp1", (size_t)C * 4, &st);
  P.hit = (float4*)ctx->buf("hit", (size_t)C * 16, &st);
  P.so = (float4*)ctx->buf("so", (size_t)C * slots * 16, &st);
  P.sd = (float4*)ctx->buf("sd", (size_t)C * slots * 16, &st);
  P.smeta = (uint2*)ctx->buf("smeta", (size_t)C * slots * 8, &st);
  P.counters = (uint32_t*)ctx->buf("counters", NUM_COUNTERS * 4, &st);
  P.stats = (unsigned long long*)ctx->buf("stats", STATS_WORDS * 8, &st);
  P.spill = (uint32_t*)ctx->buf("spill", (size_t)maxgrid * s->cfg.block * s->stack_cap * 4, &st);
  // Bounce b's shadow rays and pending NEE use set b % 2: the odd bounces' set here.  Its
  // shadow-ray tests run on stream_c[b % 2], beside the next bounces, each with a
  // traversal spill area of its own; shade adds bounce b's visible NEE at bounce b + 2.
  P.nee_lag = 2;
  P.pend_b = (float4*)ctx->buf("pend_b", (size_t)C * slots * 16, &st);
  float4* so_b = (float4*)ctx->buf("so_b", (size_t)C * slots * 16, &st);
  float4* sd_b = (float4*)ctx->buf("sd_b", (size_t)C * slots * 16, &st);
  uint2* smeta_b = (uint2*)ctx->buf("smeta_b", (size_t)C * slots * 8, &st);
  uint32_t* spill_c[2] = {(uint32_t*)ctx->buf("spill_c", (size_t)cc.grid * s->cfg.block * s->stack_cap * 4, &st),
                          (uint32_t*)ctx->buf("spill_d", (size_t)cc.grid * s->cfg.block * s->stack_cap * 4, &st)};
  hipStream_t stream_c[2] = {ctx->second_stream(), ctx->third_stream()};
  if (!stream_c[0] || !stream_c[1]) return fail(DGPU_EDEVICE, "stream creation failed");
  P.stack_cap = s->stack_cap;
  bool dev = o->flags & DGPU_BUF_DEVICE;
  if (dev) {
    P.film_rgb = film_rgb_sum;
    P.film_w = film_weight;
  } else {
    P.film_rgb = (float*)ctx->buf("film_rgb", (size_t)npix * 12, &st);
    P.film_w = (float*)ctx->buf("film_w", (size_t)npix * 4, &st);
  }
  if (st != DGPU_OK) return st;
  HIP_TRY(hipMemsetAsync(P.stats, 0, STATS_WORDS * 8, stream));
  if (!dev) {
    HIP_TRY(hipMemsetAsync(P.film_rgb, 0, (size_t)npix * 12, stream));
    HIP_TRY(hipMemsetAsync(P.film_w, 0, (size_t)npix * 4, stream));
  }
  const bool profile = (o->flags & DGPU_RENDER_PROFILE) != 0;
  std::vector<std::pair<int, size_t>> marks;  // (kernel kind, first event of the pair)
  size_t next_ev = 0;
  auto timed = [&](int kind, hipStream_t on, auto&& launch) -> hipError_t {
    if (!profile) return launch();
    hipEvent_t a = ctx->pool_event(next_ev), b = ctx->pool_event(next_ev + 1);
    if (!a || !b) return hipErrorOutOfMemory;
    hipError_t e = hipEventRecord(a, on);
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(b, on);
    marks.push_back({kind, next_ev});
    next_ev += 2;
    return e;
  };
  // Bounce b's shadow rays (connect, on stream_c[b % 2]) need only bounce b's shade, as do
  // the next bounce's extension rays (extend, on stream).  Shade b + 2, which adds bounce
  // b's visible NEE and then reuses set b % 2, waits for that connect: a bounce's
  // shadow-ray test runs beside the next bounce's extension and shade, and next to the
  // neighbouring bounces' tests, which overlaps the latency-bound ends of the launches.
  hipEvent_t ev_shade[2] = {ctx->sync_event(0), ctx->sync_event(1)};
  hipEvent_t ev_conn[2] = {ctx->sync_event(2), ctx->sync_event(3)};
  if (!ev_shade[0] || !ev_shade[1] || !ev_conn[0] || !ev_conn[1]) return fail(DGPU_EDEVICE, "event creation failed");
  WaveParams Pb[2], Pc[2];
  HIP_TRY(hipEventRecord(ctx->ev0, stream));
  bool cancelled = false;
  bool used[2] = {false, false};
  for (uint64_t g0 = 0; g0 < total && !cancelled; g0 += C) {
    P.g0 = g0;
    P.n_chunk = (uint32_t)std::min<uint64_t>(C, total - g0);
    P.q0 = (uint32_t)(g0 % npix);
    P.s0 = o->sample_begin + (uint32_t)(g0 / npix);
    Pb[0] = P;
    Pb[1] = P;
    Pb[1].pend = P.pend_b; Pb[1].so = so_b; Pb[1].sd = sd_b; Pb[1].smeta = smeta_b;
    for (int k = 0; k < 2; ++k) {
      Pc[k] = Pb[k];
      Pc[k].spill = spill_c[k];
    }
    HIP_TRY(timed(3, stream, [&] { return launch_raygen(P, P.n_chunk, false, stream); }));
    for (uint32_t b = 0; b <= o->max_depth; ++b) {
      if (cancel && *cancel) { cancelled = true; break; }
      const uint32_t k = b & 1u;
      HIP_TRY(timed(0, stream, [&] { return launch_extend(P, ce, b, false, stream); }));
      if (used[k]) HIP_TRY(hipStreamWaitEvent(stream, ev_conn[k], 0));
      HIP_TRY(timed(1, stream, [&] { return launch_shade(Pb[k], P.n_chunk, b, false, stream, shade_grid); }));
      HIP_TRY(hipEventRecord(ev_shade[k], stream));
      HIP_TRY(hipStreamWaitEvent(stream_c[k], ev_shade[k], 0));
      HIP_TRY(timed(2, stream_c[k], [&] { return launch_connect(Pc[k], cc, b, false, stream_c[k]); }));
      HIP_TRY(hipEventRecord(ev_conn[k], stream_c[k]));
      used[k] = true;
    }
    // every connect launched (after a cancel too) before what follows
    for (int k = 0; k < 2; ++k)
      if (used[k]) HIP_TRY(hipStreamWaitEvent(stream, ev_conn[k], 0));
    if (!cancelled) HIP_TRY(timed(3, stream, [&] { return launch_finalize(P, P.n_chunk, false, stream); }));
  }
  HIP_TRY(hipEventRecord(ctx->ev1, stream));
  unsigned long long hstats[STATS_WORDS] = {};
  HIP_TRY(hipMemcpyAsync(hstats, P.stats, sizeof hstats, hipMemcpyDeviceToHost, stream));
  std::vector<float> hrgb, hw;
  if (!dev && !cancelled) {
    hrgb.resize((size_t)npix * 3);
    hw.resize(npix);
    HIP_TRY(hipMemcpyAsync(hrgb.data(), P.film_rgb, (size_t)npix * 12, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(hw.data(), P.film_w, (size_t)npix * 4, hipMemcpyDeviceToHost, stream));
  }
  HIP_TRY(hipStreamSynchronize(stream));
  if (cancelled) return fail(DGPU_ECANCELLED, "render cancelled");
  if (!dev) {
    for (size_t i = 0; i < (size_t)npix * 3; ++i) film_rgb_sum[i] += hrgb[i];
    for (size_t i = 0; i < npix; ++i) film_weight[i] += hw[i];
  }
  if (out_stats) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    out_stats->paths = hstats[0];
    out_stats->extension_rays = hstats[1];
    out_stats->shadow_rays = hstats[2];
    out_stats->dropped_samples = hstats[3];
    for (int b = 0; b < 16; ++b) {
      out_stats->bounce_extension_rays[b] = hstats[4 + b];
      out_stats->bounce_shadow_rays[b] = hstats[20 + b];
    }
    out_stats->kernel_ms = ms;
    double* kms[4] = {&out_stats->extend_ms, &out_stats->shade_ms, &out_stats->connect_ms, &out_stats->other_ms};
    uint32_t* kn[4] = {&out_stats->extend_launches, &out_stats->shade_launches, &out_stats->connect_launches,
                       &out_stats->other_launches};
    for (auto& m : marks) {
      float e = 0.f;
      HIP_TRY(hipEventElapsedTime(&e, ctx->ev_pool[m.second], ctx->ev_pool[m.second + 1]));
      *kms[m.first] += e;
      *kn[m.first] += 1;
    }
  }
  return DGPU_OK;
}

// LightRenderer::render (light_renderer.cpp:13-45): x_res * y_res light paths per
// iteration, paths of min_depth + 2 .. max_depth + 2 vertices (lua_render.cpp:146-150).
dgpu_status dgpu_render_light(dgpu_scene* s, const dgpu_render_opts* o, float* film_splat_rgb,
                              dgpu_stats* out_stats, const volatile int* cancel) {
  if (!s || !o) return fail(DGPU_EINVAL, "null scene or options");
  if (!film_splat_rgb) return fail(DGPU_EINVAL, "null film buffer");
  if (o->x_res == 0 || o->y_res == 0) return fail(DGPU_EINVAL, "empty film");
  if (o->sample_end < o->sample_begin) return fail(DGPU_EINVAL, "sample_end < sample_begin");
  if (o->max_depth + 2 >= MAX_BOUNCES) return fail(DGPU_EINVAL, "max_depth must be < %u", MAX_BOUNCES - 2);
  if (o->min_depth > o->max_depth) return fail(DGPU_EINVAL, "min_depth > max_depth");
  if ((uint64_t)o->x_res * o->y_res > 0xFFFFFFFFull || o->x_res >= (1u << 28)) return fail(DGPU_EINVAL, "film too large");
  if (out_stats) memset(out_stats, 0, sizeof *out_stats);
  dgpu_ctx* ctx = s->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t stream = ctx->stream;
  dgpu_status st = DGPU_OK;
  const uint32_t npix = o->x_res * o->y_res;
  const uint64_t total = (uint64_t)npix * (o->sample_end - o->sample_begin);
  uint64_t tile = s->queue_tile ? s->queue_tile : ctx->default_tile(2);
  uint32_t C = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(total ? total : 1, tile), 0x7FFFFFFFull);
  const uint32_t grid = ctx->shade_blocks(false, false, false);

  KernelCfg ce = s->cfg, cc = s->cfg;
  ce.grid = ctx->grid_for(2, ce);
  cc.grid = ctx->grid_for(3, cc);
  uint32_t maxgrid = std::max(ce.grid, cc.grid);
  WaveParams P;
  memset(&P, 0, sizeof P);
  P.S = s->S;
  P.xr = o->x_res; P.yr = o->y_res; P.npix = npix;
  P.min_len = o->min_depth + 2; P.max_len = o->max_depth + 2;
  P.slots = 1;
  dgpu_rng_key(o->seed, o->scene_id, &P.k0, &P.k1);
  P.sample_begin = o->sample_begin;
  // one camera connection per path and stage: the shadow queue and its pending
  // records need C entries
  P.thr = (float4*)ctx->buf("thr", (size_t)C * 16, &st);
  P.rad = (float4*)ctx->buf("rad", (size_t)C * 16, &st);
  P.pend = (float4*)ctx->buf("pend", (size_t)C * 16, &st);
  P.qo[0] = (float4*)ctx->buf("qo0", (size_t)C * 16, &st);
  P.qo[1] = (float4*)ctx->buf("qo1", (size_t)C * 16, &st);
  P.qd[0] = (float4*)ctx->buf("qd0", (size_t)C * 16, &st);
  P.qd[1] = (float4*)ctx->buf("qd1", (size_t)C * 16, &st);
  P.qpath[0] = (uint32_t*)ctx->buf("qp0", (size_t)C * 4, &st);
  P.qpath[1] = (uint32_t*)ctx->buf("qp1", (size_t)C * 4, &st);
  P.hit = (float4*)ctx->buf("hit", (size_t)C * 16, &st);
  P.so = (float4*)ctx->buf("so", (size_t)C * 16, &st);
  P.sd = (float4*)ctx->buf("sd", (size_t)C * 16, &st);
  P.smeta = (uint2*)ctx->buf("smeta", (size_t)C * 8, &st);
  P.fpos = (float2*)ctx->buf("fpos", (size_t)C * 8, &st);
  P.counters = (uint32_t*)ctx->buf("counters", NUM_COUNTERS * 4, &st);
  P.stats = (unsigned long long*)ctx->buf("stats", STATS_WORDS * 8, &st);
  P.spill = (uint32_t*)ctx->buf("spill", (size_t)maxgrid * s->cfg.block * s->stack_cap * 4, &st);
  // Stage s's camera connections use the queue set s % 2 and run on stream_c[s % 2]: a
  // stage's connections have until the stage after next to finish, and consecutive
  // stages' connections (few, long rays: latency-bound launches) run side by side.
  WaveParams Ps[2] = {P, P};
  Ps[1].so = (float4*)ctx->buf("so_b", (size_t)C * 16, &st);
  Ps[1].sd = (float4*)ctx->buf("sd_b", (size_t)C * 16, &st);
  Ps[1].smeta = (uint2*)ctx->buf("smeta_b", (size_t)C * 8, &st);
  Ps[1].pend = (float4*)ctx->buf("pend_b", (size_t)C * 16, &st);
  Ps[1].fpos = (float2*)ctx->buf("fpos_b", (size_t)C * 8, &st);
  uint32_t* spill_c[2] = {(uint32_t*)ctx->buf("spill_c", (size_t)cc.grid * s->cfg.block * s->stack_cap * 4, &st),
                          (uint32_t*)ctx->buf("spill_d", (size_t)cc.grid * s->cfg.block * s->stack_cap * 4, &st)};
  hipStream_t stream_c[2] = {ctx->second_stream(), ctx->third_stream()};
  if (!stream_c[0] || !stream_c[1]) return fail(DGPU_EDEVICE, "stream creation failed");
  P.stack_cap = s->stack_cap;
  bool dev = o->flags & DGPU_BUF_DEVICE;
  P.film_rgb = dev ? film_splat_rgb : (float*)ctx->buf("film_rgb", (size_t)npix * 12, &st);
  if (st != DGPU_OK) return st;
  HIP_TRY(hipMemsetAsync(P.stats, 0, STATS_WORDS * 8, stream));
  if (!dev) HIP_TRY(hipMemsetAsync(P.film_rgb, 0, (size_t)npix * 12, stream));
  const bool profile = (o->flags & DGPU_RENDER_PROFILE) != 0;
  std::vector<std::pair<int, size_t>> marks;
  size_t next_ev = 0;
  auto timed = [&](int kind, hipStream_t on, auto&& launch) -> hipError_t {
    if (!profile) return launch();
    hipEvent_t a = ctx->pool_event(next_ev), b = ctx->pool_event(next_ev + 1);
    if (!a || !b) return hipErrorOutOfMemory;
    hipError_t e = hipEventRecord(a, on);
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(b, on);
    marks.push_back({kind, next_ev});
    next_ev += 2;
    return e;
  };
  // The next stage's extension (extend, on stream) needs only the stage's shade, as do
  // the stage's camera connections (connect, then splat): they run side by side.  A
  // shade that writes stage s's connections waits for the splat of stage s - 2, the last
  // user of the queue set s % 2.
  hipEvent_t ev_shade[2] = {ctx->sync_event(0), ctx->sync_event(1)};
  hipEvent_t ev_splat[2] = {ctx->sync_event(2), ctx->sync_event(3)};
  if (!ev_shade[0] || !ev_shade[1] || !ev_splat[0] || !ev_splat[1])
    return fail(DGPU_EDEVICE, "event creation failed");
  WaveParams Pc[2] = {P, P};
  HIP_TRY(hipEventRecord(ctx->ev0, stream));
  bool cancelled = false;
  bool used[2] = {false, false};  // a set's splat recorded in this render
  auto connect_stage = [&](uint32_t stage) -> hipError_t {
    const uint32_t k = stage & 1u;
    hipError_t e = hipEventRecord(ev_shade[k], stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(stream_c[k], ev_shade[k], 0);
    if (e == hipSuccess) e = timed(2, stream_c[k], [&] { return launch_connect(Pc[k], cc, stage, false, stream_c[k]); });
    if (e == hipSuccess) e = timed(3, stream_c[k], [&] { return launch_lt_splat(Pc[k], stage, stream_c[k], grid); });
    if (e == hipSuccess) e = hipEventRecord(ev_splat[k], stream_c[k]);
    used[k] = true;
    return e;
  };
  for (uint64_t g0 = 0; g0 < total && !cancelled; g0 += C) {
    for (int k = 0; k < 2; ++k) {
      WaveParams q = P;
      q.so = Ps[k].so; q.sd = Ps[k].sd; q.smeta = Ps[k].smeta; q.pend = Ps[k].pend; q.fpos = Ps[k].fpos;
      q.g0 = g0;
      q.n_chunk = (uint32_t)std::min<uint64_t>(C, total - g0);
      Ps[k] = q;
      Pc[k] = q;
      Pc[k].spill = spill_c[k];
    }
    P.g0 = Ps[0].g0;
    P.n_chunk = Ps[0].n_chunk;
    // a new chunk zeroes the counters: every connection of the last one first
    for (int k = 0; k < 2; ++k)
      if (used[k]) HIP_TRY(hipStreamWaitEvent(stream, ev_splat[k], 0));
    HIP_TRY(launch_zero(P.counters, NUM_COUNTERS, stream));
    HIP_TRY(timed(1, stream, [&] { return launch_lt_start(Ps[0], stream, grid); }));
    HIP_TRY(connect_stage(0));
    for (uint32_t b = 0; b + 3u <= P.max_len; ++b) {
      if (cancel && *cancel) { cancelled = true; break; }
      const uint32_t k = (b + 1u) & 1u;  // the set of stage b + 1, which this shade writes
      HIP_TRY(timed(0, stream, [&] { return launch_extend(P, ce, b, false, stream); }));
      if (used[k]) HIP_TRY(hipStreamWaitEvent(stream, ev_splat[k], 0));
      HIP_TRY(timed(1, stream, [&] { return launch_lt_shade(Ps[k], b, stream, grid); }));
      HIP_TRY(connect_stage(b + 1));
    }
    for (int k = 0; k < 2; ++k)
      if (used[k]) HIP_TRY(hipStreamWaitEvent(stream, ev_splat[k], 0));
    HIP_TRY(launch_lt_stats(P, stream));
  }
  HIP_TRY(hipEventRecord(ctx->ev1, stream));
  unsigned long long hstats[STATS_WORDS] = {};
  HIP_TRY(hipMemcpyAsync(hstats, P.stats, sizeof hstats, hipMemcpyDeviceToHost, stream));
  std::vector<float> hrgb;
  if (!dev && !cancelled) {
    hrgb.resize((size_t)npix * 3);
    HIP_TRY(hipMemcpyAsync(hrgb.data(), P.film_rgb, (size_t)npix * 12, hipMemcpyDeviceToHost, stream));
  }
  HIP_TRY(hipStreamSynchronize(stream));
  if (cancelled) return fail(DGPU_ECANCELLED, "render cancelled");
  if (!dev)
    for (size_t i = 0; i < (size_t)npix * 3; ++i) film_splat_rgb[i] += hrgb[i];
  if (out_stats) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    out_stats->paths = hstats[0];
    out_stats->extension_rays = hstats[1];
    out_stats->shadow_rays = hstats[2];
    out_stats->dropped_samples = hstats[3];
    for (int b = 0; b < 16; ++b) {
      out_stats->bounce_extension_rays[b] = hstats[4 + b];
      out_stats->bounce_shadow_rays[b] = hstats[20 + b];
    }
    out_stats->kernel_ms = ms;
    double* kms[4] = {&out_stats->extend_ms, &out_stats->shade_ms, &out_stats->connect_ms, &out_stats->other_ms};
    uint32_t* kn[4] = {&out_stats->extend_launches, &out_stats->shade_launches, &out_stats->connect_launches,
                       &out_stats->other_launches};
    for (auto& m : marks) {
      float e = 0.f;
      HIP_TRY(hipEventElapsedTime(&e, ctx->ev_pool[m.second], ctx->ev_pool[m.second + 1]));
      *kms[m.first] += e;
      *kn[m.first] += 1;
    }
  }
  return DGPU_OK;
}

dgpu_status dgpu_scene_update_texture(dgpu_scene* s, uint32_t tex, const float* rgb) {
  if (!s || !rgb) return fail(DGPU_EINVAL, "null scene or texels");
  if (tex >= s->tex_info_host.size()) return fail(DGPU_EINVAL, "texture %u out of range", tex);
  uint4 info = s->tex_info_host[tex];
  size_t n = (size_t)info.x * info.y;
  std::vector<float4> t(n);
  for (size_t i = 0; i < n; ++i) t[i] = make_float4(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], 0.f);
  dgpu_ctx* ctx = s->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMemcpy((void*)(s->S.tex_data + info.z), t.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  return DGPU_OK;
}

dgpu_status dgpu_scene_update_material(dgpu_scene* s, uint32_t material, const dgpu_material* m) {
  if (!s || !m) return fail(DGPU_EINVAL, "null scene or material");
  if (material >= s->mats_host.size()) return fail(DGPU_EINVAL, "material %u out of range", material);
  if (m->type > DGPU_MAT_ROUGH_DIELECTRIC) return fail(DGPU_EUNSUPPORTED, "unsupported material type %u", m->type);
  if (m->type == DGPU_MAT_LAMBERT && m->albedo_tex >= (int32_t)s->tex_info_host.size())
    return fail(DGPU_EINVAL, "texture %d out of range", m->albedo_tex);
  if (m->type == DGPU_MAT_GGX_CONDUCTOR && !(m->roughness > 0.f))
    return fail(DGPU_EINVAL, "GGX roughness must be > 0");
  if (m->type == DGPU_MAT_ROUGH_DIELECTRIC && !(m->roughness >= 0.f && std::isfinite(m->roughness)))
    return fail(DGPU_EINVAL, "rough dielectric roughness must be finite and >= 0");
  DevMaterial dm = to_dev_material(*m);
  dgpu_ctx* ctx = s->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipMemcpy((void*)(s->S.mats + material), &dm, sizeof dm, hipMemcpyHostToDevice));
  s->mats_host[material] = dm;
  s->S.has_rough = 0;
  for (const DevMaterial& h : s->mats_host) s->S.has_rough |= h.type == DGPU_MAT_ROUGH_DIELECTRIC ? 1u : 0u;
  return DGPU_OK;
}

dgpu_status dgpu_render_bdpt(dgpu_scene* s, const dgpu_render_opts* o, uint32_t use_t1_paths, float* film_rgb,
                             float* film_w, float* film_splat_rgb, dgpu_stats* out_stats,
                             const volatile int* cancel) {
  if (!s || !o) return fail(DGPU_EINVAL, "null scene or options");
  if (!film_rgb || !film_w || !film_splat_rgb) return fail(DGPU_EINVAL, "null film buffer");
  if (o->x_res == 0 || o->y_res == 0) return fail(DGPU_EINVAL, "empty film");
  if (o->sample_end < o->sample_begin) return fail(DGPU_EINVAL, "sample_end < sample_begin");
  if (o->max_depth >= 30) return fail(DGPU_EINVAL, "max_depth must be < 30");
  if (o->min_depth > o->max_depth) return fail(DGPU_EINVAL, "min_depth > max_depth");
  if ((uint64_t)o->x_res * o->y_res > 0xFFFFFFFFull || o->x_res >= (1u << 28)) return fail(DGPU_EINVAL, "film too large");
  if (out_stats) memset(out_stats, 0, sizeof *out_stats);
  dgpu_ctx* ctx = s->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t stream = ctx->stream;
  dgpu_status st = DGPU_OK;
  const uint32_t npix = o->x_res * o->y_res;
  // BdptRenderer::render returns at once without lights (bdpt_renderer.cpp:11)
  const uint64_t total = s->S.num_lights ? (uint64_t)npix * (o->sample_end - o->sample_begin) : 0;
  const uint32_t nv = o->max_depth + 2u;
  const bool mega = o->flags & DGPU_BDPT_MEGAKERNEL;
  // strategies (s, t) with t >= 1 and s + t <= nv: the records per path of the wavefront form
  const uint32_t ns = nv * (nv + 1u) / 2u;
  // per path: two subpaths of nv vertices and the staged sample; the megakernel adds its
  // traversal spill slice.  The wavefront form adds its walk state and two extension rays
  // per bounce (both queues and the hits); its strategy stage, which runs over parts of
  // the chunk, needs per path ns records (24 B) and shadow-queue slots (40 B).
  const uint64_t per_path = mega ? 2ull * nv * BDPT_VERTEX_BYTES + 16ull + 4ull * s->stack_cap
                                 : 2ull * nv * BDPT_VERTEX_BYTES + 16ull + 64ull + 8ull + 176ull;
  const uint64_t per_strat = (uint64_t)ns * 64ull;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  const uint64_t budget = std::min<uint64_t>(mega ? (8ull << 30) : (32ull << 30), free_b / 4);
  // the wavefront form gives half of its budget to the walks, half to the strategy stage
  uint64_t cap = std::max<uint64_t>(1ull << 16, (mega ? budget : budget / 2) / per_path);
  if (s->queue_tile) cap = s->queue_tile;
  // chunk positions q0 + i below 2^32; both walks' queue words (path index | camera bit)
  // below 2^31 entries
  const uint32_t C = (uint32_t)std::min<uint64_t>(
      std::min<uint64_t>(std::min<uint64_t>(total ? total : 1, cap), 0x3FFFFFC0ull),
      std::max<uint64_t>(1, 0xFFFFFFFFull - npix));
  // paths per strategy part: record indices j * ns below 2^32
  const uint32_t M = (uint32_t)std::min<uint64_t>(
      std::min<uint64_t>(C, std::max<uint64_t>(1ull << 16, budget / 2 / per_strat)), 0xFFFFFFFFull / ns);
  WaveParams P;
  memset(&P, 0, sizeof P);
  P.S = s->S;
  P.xr = o->x_res; P.yr = o->y_res; P.npix = npix;
  P.ord = make_path_order(o->x_res, o->y_res);
  P.min_depth = o->min_depth; P.max_depth = o->max_depth;
  dgpu_rng_key(o->seed, o->scene_id, &P.k0, &P.k1);
  P.sample_begin = o->sample_begin;
  BdptArgsHost A;
  memset(&A, 0, sizeof A);
  A.n = C;
  A.nv = nv;
  A.ns = ns;
  A.use_t1 = use_t1_paths ? 1u : 0u;
  A.lv = (float4*)ctx->buf("bd_lv", (size_t)C * nv * BDPT_VERTEX_BYTES, &st);
  A.cv = (float4*)ctx->buf("bd_cv", (size_t)C * nv * BDPT_VERTEX_BYTES, &st);
  P.rad = (float4*)ctx->buf("rad", (size_t)C * 16, &st);
  P.stats = (unsigned long long*)ctx->buf("stats", STATS_WORDS * 8, &st);
  P.stack_cap = s->stack_cap;
  KernelCfg ce = s->cfg, cc = s->cfg;
  uint32_t shade_grid = 0;
  if (mega) {
    const size_t threads = ((size_t)C + 63) / 64 * 64;
    P.spill = (uint32_t*)ctx->buf("bd_spill", threads * s->stack_cap * 4, &st);
  } else {
    ce.grid = ctx->grid_for(2, ce);
    cc.grid = ctx->grid_for(3, cc);
    shade_grid = ctx->shade_blocks(false, false, false);
    P.spill = (uint32_t*)ctx->buf("spill", (size_t)std::max(ce.grid, cc.grid) * s->cfg.block * s->stack_cap * 4, &st);
    const size_t q = (size_t)2 * C;  // a light and a camera ray per path and bounce
    P.qo[0] = (float4*)ctx->buf("qo0", q * 16, &st);
    P.qo[1] = (float4*)ctx->buf("qo1", q * 16, &st);
    P.qd[0] = (float4*)ctx->buf("qd0", q * 16, &st);
    P.qd[1] = (float4*)ctx->buf("qd1", q * 16, &st);
    P.qpath[0] = (uint32_t*)ctx->buf("qp0", q * 4, &st);
    P.qpath[1] = (uint32_t*)ctx->buf("qp1", q * 4, &st);
    P.hit = (float4*)ctx->buf("hit", q * 16, &st);
    P.so = (float4*)ctx->buf("so", (size_t)M * ns * 16, &st);
    P.sd = (float4*)ctx->buf("sd", (size_t)M * ns * 16, &st);
    P.smeta = (uint2*)ctx->buf("smeta", (size_t)M * ns * 8, &st);
    P.counters = (uint32_t*)ctx->buf("counters", NUM_COUNTERS * 4, &st);
    A.w0 = (float4*)ctx->buf("bd_w0", q * 16, &st);  // light walks, then camera walks
    A.w1 = (uint4*)ctx->buf("bd_w1", q * 16, &st);
    A.nl = (uint32_t*)ctx->buf("bd_nl", (size_t)C * 4, &st);
    A.nc = (uint32_t*)ctx->buf("bd_nc", (size_t)C * 4, &st);
    A.rec = (float4*)ctx->buf("bd_rec", (size_t)M * ns * 16, &st);
    A.rec_fp = (float2*)ctx->buf("bd_rec_fp", (size_t)M * ns * 8, &st);
    P.pend = A.rec;  // k_connect clears an occluded ray's record (its .w)
  }
  const bool dev = o->flags & DGPU_BUF_DEVICE;
  P.film_rgb = dev ? film_rgb : (float*)ctx->buf("film_rgb", (size_t)npix * 12, &st);
  P.film_w = dev ? film_w : (float*)ctx->buf("film_w", (size_t)npix * 4, &st);
  A.splat = dev ? film_splat_rgb : (float*)ctx->buf("film_splat", (size_t)npix * 12, &st);
  if (st != DGPU_OK) return st;
  HIP_TRY(hipMemsetAsync(P.stats, 0, STATS_WORDS * 8, stream));
  if (!dev) {
    HIP_TRY(hipMemsetAsync(P.film_rgb, 0, (size_t)npix * 12, stream));
    HIP_TRY(hipMemsetAsync(P.film_w, 0, (size_t)npix * 4, stream));
    HIP_TRY(hipMemsetAsync(A.splat, 0, (size_t)npix * 12, stream));
  }
  // the wavefront form's strategy kernels run on three streams
  const hipStream_t side[2] = {ctx->second_stream(), ctx->third_stream()};
  const hipEvent_t sev[3] = {ctx->sync_event(0), ctx->sync_event(1), ctx->sync_event(2)};
  if (!side[0] || !side[1] || !sev[0] || !sev[1] || !sev[2]) return fail(DGPU_EDEVICE, "stream/event creation failed");
  HIP_TRY(hipEventRecord(ctx->ev0, stream));
  bool cancelled = false;
  for (uint64_t g0 = 0; g0 < total; g0 += C) {
    if (cancel && *cancel) { cancelled = true; break; }
    P.g0 = g0;
    P.n_chunk = (uint32_t)std::min<uint64_t>(C, total - g0);
    P.q0 = (uint32_t)(g0 % npix);
    P.s0 = o->sample_begin + (uint32_t)(g0 / npix);
    if (mega) HIP_TRY(launch_bdpt(P, A, s->S.width, stream));
    else HIP_TRY(launch_bdpt_wave(P, A, ce, cc, shade_grid, M, stream, side, sev));
  }
  HIP_TRY(hipEventRecord(ctx->ev1, stream));
  unsigned long long hstats[STATS_WORDS] = {};
  HIP_TRY(hipMemcpyAsync(hstats, P.stats, sizeof hstats, hipMemcpyDeviceToHost, stream));
  std::vector<float> hrgb, hw, hsp;
  if (!dev && !cancelled) {
    hrgb.resize((size_t)npix * 3);
    hw.resize(npix);
    hsp.resize((size_t)npix * 3);
    HIP_TRY(hipMemcpyAsync(hrgb.data(), P.film_rgb, (size_t)npix * 12, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(hw.data(), P.film_w, (size_t)npix * 4, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(hsp.data(), A.splat, (size_t)npix * 12, hipMemcpyDeviceToHost, stream));
  }
  HIP_TRY(hipStreamSynchronize(stream));
  if (cancelled) return fail(DGPU_ECANCELLED, "render cancelled");
  if (!dev) {
    for (size_t k = 0; k < (size_t)npix * 3; ++k) {
      film_rgb[k] += hrgb[k];
      film_splat_rgb[k] += hsp[k];
    }
    for (size_t k = 0; k < npix; ++k) film_w[k] += hw[k];
  }
  if (out_stats) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    out_stats->paths = hstats[0];
    out_stats->extension_rays = hstats[1];
    out_stats->shadow_rays = hstats[2];
    out_stats->dropped_samples = hstats[3];
    out_stats->kernel_ms = ms;
  }
  return DGPU_OK;
}

dgpu_status dgpu_adjoint(dgpu_scene* s, const dgpu_render_opts* o, const float* dldi, float* d_tex,
                         float* d_rough, float* image_out, dgpu_stats* out_stats) {
  if (!s || !o || !dldi) return fail(DGPU_EINVAL, "null scene, options or dL/dI");
  if (o->x_res == 0 || o->y_res == 0) return fail(DGPU_EINVAL, "empty film");
  if (o->sample_end <= o->sample_begin) return fail(DGPU_EINVAL, "empty sample range");
  if (o->max_depth + 2 >= MAX_BOUNCES) return fail(DGPU_EINVAL, "max_depth must be < %u", MAX_BOUNCES - 2);
  if (o->direct_strategy > DGPU_DIRECT_LIGHT) return fail(DGPU_EINVAL, "bad direct_strategy");
  if ((uint64_t)o->x_res * o->y_res > 0xFFFFFFFFull || o->x_res >= (1u << 28)) return fail(DGPU_EINVAL, "film too large");
  if (o->sample_all_lights && s->S.num_lights > MAX_SLOT_LIGHTS)
    return fail(DGPU_EUNSUPPORTED, "sample_all_lights supports at most %u lights", MAX_SLOT_LIGHTS);
  if (d_tex && s->tex_info_host.empty()) return fail(DGPU_EINVAL, "scene has no texture 0");
  if (out_stats) memset(out_stats, 0, sizeof *out_stats);
  dgpu_ctx* ctx = s->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t stream = ctx->stream;
  dgpu_status st = DGPU_OK;
  const uint32_t npix = o->x_res * o->y_res;
  const uint32_t spp = o->sample_end - o->sample_begin;
  const uint64_t total = (uint64_t)npix * spp;
  const size_t texels = s->tex_info_host.empty() ? 0 : (size_t)s->tex_info_host[0].x * s->tex_info_host[0].y;
  const uint32_t nm = s->S.num_mats;
  const bool dev = o->flags & DGPU_BUF_DEVICE;
  const bool mega = o->flags & DGPU_ADJOINT_MEGAKERNEL;
  // gradient, dL/dI and image buffers on the device (the caller's, or staging)
  const float* a_dldi;
  float *a_tex, *a_rough, *a_img;
  if (dev) {
    a_dldi = dldi; a_tex = d_tex; a_rough = d_rough; a_img = image_out;
  } else {
    float* b_dldi = (float*)ctx->buf("adj_dldi", (size_t)npix * 12, &st);
    float* b_tex = d_tex ? (float*)ctx->buf("adj_gtex", texels * 12, &st) : nullptr;
    float* b_rough = d_rough ? (float*)ctx->buf("adj_grough", (size_t)nm * 4, &st) : nullptr;
    float* b_img = image_out ? (float*)ctx->buf("adj_img", (size_t)npix * 12, &st) : nullptr;
    if (st != DGPU_OK) return st;
    HIP_TRY(hipMemcpyAsync(b_dldi, dldi, (size_t)npix * 12, hipMemcpyHostToDevice, stream));
    if (b_tex) HIP_TRY(hipMemsetAsync(b_tex, 0, texels * 12, stream));
    if (b_rough) HIP_TRY(hipMemsetAsync(b_rough, 0, (size_t)nm * 4, stream));
    if (b_img) HIP_TRY(hipMemsetAsync(b_img, 0, (size_t)npix * 12, stream));
    a_dldi = b_dldi; a_tex = b_tex; a_rough = b_rough; a_img = b_img;
  }
  WaveParams P;
  memset(&P, 0, sizeof P);
  P.S = s->S;
  P.xr = o->x_res; P.yr = o->y_res; P.npix = npix;
  P.ord = make_path_order(o->x_res, o->y_res);
  P.min_depth = o->min_depth; P.max_depth = o->max_depth; P.rr_depth = o->rr_depth;
  P.only_direct = o->only_direct; P.sample_all = o->sample_all_lights; P.strategy = o->direct_strategy;
  P.rough = s->S.has_rough;
  dgpu_rng_key(o->seed, o->scene_id, &P.k0, &P.k1);
  P.sample_begin = o->sample_begin;
  if (mega) {
    // one thread per path: trace it, then replay it with the deposits (k_adjoint)
    P.slots = 2;
    uint32_t chunk = (uint32_t)std::min<uint64_t>(total, ADJ_CHUNK);
    size_t threads = ((size_t)chunk + 63) / 64 * 64;
    P.spill = (uint32_t*)ctx->buf("adj_spill", threads * s->stack_cap * 4, &st);
    if (st != DGPU_OK) return st;
    AdjArgs A;
    memset(&A, 0, sizeof A);
    A.inv_spp = 1.f / float(spp);
    A.dldi = a_dldi; A.g_tex = a_tex; A.g_rough = a_rough; A.image = a_img;
    HIP_TRY(hipEventRecord(ctx->ev0, stream));
    for (uint64_t g0 = 0; g0 < total; g0 += chunk) {
      P.g0 = g0;
      P.n_chunk = (uint32_t)std::min<uint64_t>(chunk, total - g0);
      HIP_TRY(launch_adjoint(P, A, s->S.width, stream));
    }
    HIP_TRY(hipEventRecord(ctx->ev1, stream));
  } else {
    // Wavefront replay.  Per chunk, pass 1 renders the paths (the path tracer's kernels)
    // and keeps each path's radiance; pass 2 renders the same paths again (the same
    // counter-based random numbers) with the adjoint shade, whose lobes leave gradient
    // payloads that are deposited once the radiance they carry is known: a shadow ray's
    // at the next bounce (or after the last), a continuation's at the next bounce as the
    // first-pass radiance less what the path has gathered by then.
    const uint32_t slots = o->sample_all_lights ? 2 * std::max(1u, s->S.num_lights) : 2;
    P.slots = slots;
    const uint64_t per_path = 120ull + 56ull * slots + 32ull + 16ull * slots;  // + first-pass radiance, payloads
    uint64_t tile = s->queue_tile;
    if (!tile) {
      size_t free_b = 0, total_b = 0;
      if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
      tile = std::max<uint64_t>(1ull << 20, std::min<uint64_t>(16ull << 30, free_b / 4) / per_path);
    }
    const uint32_t C = (uint32_t)std::min<uint64_t>(
        std::min<uint64_t>(total, tile),
        std::min<uint64_t>(0xFFFFFFFFull / slots, std::max<uint64_t>(1, 0xFFFFFFFFull - npix)));
    const uint32_t shade_grid = ctx->shade_blocks(false, o->sample_all_lights != 0, s->S.has_rough != 0);
    const uint32_t shade_grid_adj = ctx->shade_blocks_adjoint(o->sample_all_lights != 0, s->S.has_rough != 0);
    KernelCfg ce = s->cfg, cc = s->cfg;
    ce.grid = ctx->grid_for(2, ce);
    cc.grid = ctx->grid_for(3, cc);
    uint32_t maxgrid = std::max(ce.grid, cc.grid);
    P.thr = (float4*)ctx->buf("thr", (size_t)C * 16, &st);
    P.rad = (float4*)ctx->buf("rad", (size_t)C * 16, &st);
    P.pend = (float4*)ctx->buf("pend", (size_t)C * slots * 16, &st);
    P.qo[0] = (float4*)ctx->buf("qo0", (size_t)C * 16, &st);
    P.qo[1] = (float4*)ctx->buf("qo1", (size_t)C * 16, &st);
    P.qd[0] = (float4*)ctx->buf("qd0", (size_t)C * 16, &st);
    P.qd[1] = (float4*)ctx->buf("qd1", (size_t)C * 16, &st);
    P.qpath[0] = (uint32_t*)ctx->buf("qp0", (size_t)C * 4, &st);
    P.qpath[1] = (uint32_t*)ctx->buf("qp1", (size_t)C * 4, &st);
    P.hit = (float4*)ctx->buf("hit", (size_t)C * 16, &st);
    P.so = (float4*)ctx->buf("so", (size_t)C * slots * 16, &st);
    P.sd = (float4*)ctx->buf("sd", (size_t)C * slots * 16, &st);
    P.smeta = (uint2*)ctx->buf("smeta", (size_t)C * slots * 8, &st);
    P.counters = (uint32_t*)ctx->buf("counters", NUM_COUNTERS * 4, &st);
    P.stats = (unsigned long long*)ctx->buf("stats", STATS_WORDS * 8, &st);
    P.spill = (uint32_t*)ctx->buf("spill", (size_t)maxgrid * s->cfg.block * s->stack_cap * 4, &st);
    P.stack_cap = s->stack_cap;
    P.adj_L = (float4*)ctx->buf("adj_L", (size_t)C * 16, &st);
    P.adj_pend = (float4*)ctx->buf("adj_pend", (size_t)C * slots * 16, &st);
    P.adj_cont = (float4*)ctx->buf("adj_cont", (size_t)C * 16, &st);
    uint32_t* spill_c = (uint32_t*)ctx->buf("spill_c", (size_t)cc.grid * s->cfg.block * s->stack_cap * 4, &st);
    hipStream_t stream_c = ctx->second_stream();
    hipEvent_t ev_shade = ctx->sync_event(0), ev_conn = ctx->sync_event(1);
    if (!stream_c || !ev_shade || !ev_conn) return fail(DGPU_EDEVICE, "stream/event creation failed");
    if (st != DGPU_OK) return st;
    P.adj_dldi = a_dldi;
    P.adj_g_tex = a_tex;
    P.adj_g_rough = a_rough;
    P.adj_image = a_img;
    P.adj_inv_spp = 1.f / float(spp);
    HIP_TRY(hipMemsetAsync(P.stats, 0, STATS_WORDS * 8, stream));
    HIP_TRY(hipEventRecord(ctx->ev0, stream));
    for (uint64_t g0 = 0; g0 < total; g0 += C) {
      P.g0 = g0;
      P.n_chunk = (uint32_t)std::min<uint64_t>(C, total - g0);
      P.q0 = (uint32_t)(g0 % npix);
      P.s0 = o->sample_begin + (uint32_t)(g0 / npix);
      WaveParams Pc = P;
      Pc.spill = spill_c;
      for (int pass = 0; pass < 2; ++pass) {
        // as in dgpu_render: bounce b's connect (stream_c) beside bounce b + 1's extend;
        // shade b + 1 takes bounce b's visible NEE, so it waits for that connect
        HIP_TRY(launch_raygen(P, P.n_chunk, false, stream));
        for (uint32_t b = 0; b <= o->max_depth; ++b) {
          HIP_TRY(launch_extend(P, ce, b, false, stream));
          if (b > 0) HIP_TRY(hipStreamWaitEvent(stream, ev_conn, 0));
          if (pass == 0) HIP_TRY(launch_shade(P, P.n_chunk, b, false, stream, shade_grid));
          else HIP_TRY(launch_shade_adjoint(P, P.n_chunk, b, stream, shade_grid_adj));
          HIP_TRY(hipEventRecord(ev_shade, stream));
          HIP_TRY(hipStreamWaitEvent(stream_c, ev_shade, 0));
          HIP_TRY(launch_connect(Pc, cc, b, false, stream_c));
          HIP_TRY(hipEventRecord(ev_conn, stream_c));
        }
        HIP_TRY(hipStreamWaitEvent(stream, ev_conn, 0));
        if (pass == 0) HIP_TRY(launch_adj_total(P, stream));
        else HIP_TRY(launch_adj_final(P, stream));
      }
    }
    HIP_TRY(hipEventRecord(ctx->ev1, stream));
  }
  std::vector<float> htex, hrough, himg;
  if (!dev) {
    if (d_tex) { htex.resize(texels * 3); HIP_TRY(hipMemcpyAsync(htex.data(), a_tex, texels * 12, hipMemcpyDeviceToHost, stream)); }
    if (d_rough) { hrough.resize(nm); HIP_TRY(hipMemcpyAsync(hrough.data(), a_rough, (size_t)nm * 4, hipMemcpyDeviceToHost, stream)); }
    if (image_out) { himg.resize((size_t)npix * 3); HIP_TRY(hipMemcpyAsync(himg.data(), a_img, (size_t)npix * 12, hipMemcpyDeviceToHost, stream)); }
  }
  HIP_TRY(hipStreamSynchronize(stream));
  if (!dev) {
    for (size_t i = 0; i < htex.size(); ++i) d_tex[i] += htex[i];
    for (size_t i = 0; i < hrough.size(); ++i) d_rough[i] += hrough[i];
    for (size_t i = 0; i < himg.size(); ++i) image_out[i] += himg[i];
  }
  if (out_stats) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    out_stats->paths = total;
    out_stats->kernel_ms = ms;
  }
  return DGPU_OK;
}

dgpu_status dgpu_render_batch(dgpu_ctx* ctx, const dgpu_batch_item* items, uint32_t n, uint32_t flags,
                              dgpu_stats* out_stats) {
  if (!ctx || !items) return fail(DGPU_EINVAL, "null context or items");
  if (n == 0 || n > MAX_BATCH) return fail(DGPU_EINVAL, "batch size must be in [1, %u]", MAX_BATCH);
  const dgpu_render_opts& o0 = items[0].opts;
  const dgpu_scene* s0 = items[0].scene;
  if (!s0) return fail(DGPU_EINVAL, "item 0: null scene");
  for (uint32_t i = 0; i < n; ++i) {
    const dgpu_batch_item& it = items[i];
    const dgpu_render_opts& o = it.opts;
    if (!it.scene || it.scene->ctx != ctx) return fail(DGPU_EINVAL, "item %u: scene of another context", i);
    if (!it.film_rgb_sum || !it.film_weight) return fail(DGPU_EINVAL, "item %u: null film", i);
    if (o.x_res == 0 || o.y_res == 0) return fail(DGPU_EINVAL, "item %u: empty film", i);
    if ((uint64_t)o.x_res * o.y_res > (1u << 30) || o.x_res >= (1u << 28))
      return fail(DGPU_EINVAL, "item %u: film too large", i);
    if (o.sample_end < o.sample_begin) return fail(DGPU_EINVAL, "item %u: sample_end < sample_begin", i);
    if (o.min_depth != o0.min_depth || o.max_depth != o0.max_depth || o.rr_depth != o0.rr_depth ||
        o.only_direct != o0.only_direct || o.direct_strategy != o0.direct_strategy)
      return fail(DGPU_EINVAL, "item %u: depth / strategy options differ from item 0", i);
    if (o.sample_all_lights) return fail(DGPU_EUNSUPPORTED, "batches do not support sample_all_lights");
  }
  if (o0.max_depth + 2 >= MAX_BOUNCES) return fail(DGPU_EINVAL, "max_depth must be < %u", MAX_BOUNCES - 2);
  if (o0.direct_strategy > DGPU_DIRECT_LIGHT) return fail(DGPU_EINVAL, "bad direct_strategy");
  if (out_stats) memset(out_stats, 0, sizeof *out_stats);
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t stream = ctx->stream;
  dgpu_status st = DGPU_OK;
  const bool dev = flags & DGPU_BUF_DEVICE;
  const bool use_graph = !(flags & DGPU_BATCH_NO_GRAPH);

  // scene table and (for host films) one device film region
  std::vector<BatchEntry> tab(n);
  std::vector<uint64_t> paths(n);
  size_t film_floats = 0;
  uint32_t stack_cap = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const dgpu_render_opts& o = items[i].opts;
    BatchEntry& E = tab[i];
    memset(&E, 0, sizeof E);
    E.S = items[i].scene->S;
    E.xr = o.x_res; E.yr = o.y_res; E.npix = o.x_res * o.y_res;
    E.ord = make_path_order(o.x_res, o.y_res);
    E.sample_begin = o.sample_begin;
    dgpu_rng_key(o.seed, o.scene_id, &E.k0, &E.k1);
    paths[i] = (uint64_t)E.npix * (o.sample_end - o.sample_begin);
    film_floats += 4 * (size_t)E.npix;
    stack_cap = std::max(stack_cap, items[i].scene->stack_cap);
  }
  if (dev) {
    // The finalize kernel reads, adds and writes each film pixel without atomics, and
    // two items of one chunk are finalized by different threads: device films (and an
    // item's colour and weight arrays) must not share memory.  Host films do not
    // matter: each item gets its own staging region, added to the host film at the end.
    struct Span { uintptr_t lo, hi; uint32_t item; };
    std::vector<Span> spans;
    spans.reserve(2 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
      uintptr_t r = (uintptr_t)items[i].film_rgb_sum, w = (uintptr_t)items[i].film_weight;
      spans.push_back({r, r + 12 * (uintptr_t)tab[i].npix, i});
      spans.push_back({w, w + 4 * (uintptr_t)tab[i].npix, i});
    }
    std::sort(spans.begin(), spans.end(), [](const Span& a, const Span& b) { return a.lo < b.lo; });
    for (size_t k = 1; k < spans.size(); ++k)
      if (spans[k].lo < spans[k - 1].hi)
        return fail(DGPU_EINVAL, "items %u and %u: device films overlap (each film must be its own memory)",
                    std::min(spans[k - 1].item, spans[k].item), std::max(spans[k - 1].item, spans[k].item));
  }
  float* dfilm = nullptr;
  if (!dev) {
    dfilm = (float*)ctx->buf("batch_film", film_floats * 4, &st);
    if (st != DGPU_OK) return st;
    HIP_TRY(hipMemsetAsync(dfilm, 0, film_floats * 4, stream));
  }
  size_t off = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (dev) {
      tab[i].film_rgb = items[i].film_rgb_sum;
      tab[i].film_w = items[i].film_weight;
    } else {
      tab[i].film_rgb = dfilm + off;
      tab[i].film_w = dfilm + off + 3 * (size_t)tab[i].npix;
      off += 4 * (size_t)tab[i].npix;
    }
  }

  // Items are grouped by their scene's kernel knobs (BVH width, LDS stack entries,
  // trace workgroup); each group's paths are cut into chunks of its own, so a chunk is
  // traced by one kernel variant.  Chunks run over the scene-major concatenation of the
  // group's paths.  Queue words hold chunk path indices (NEE slot 2 * path + k), and the
  // kernels find a path's scene by a search over the chunk's segment starts, so a chunk
  // can hold the whole group when memory allows.  Larger chunks mean fewer bounce
  // launches and shorter tails (64 scenes: 2^23-path chunks rendered the batch workload
  // in 319 ms, 2^25 in 227 ms, 2^26 in 212 ms, and chunks of a third of the batch, 110M
  // paths, in 202 ms; 2^28 and one 331M-path chunk took 207 ms).  A group's chunks are
  // cut to equal sizes.
  uint64_t total = 0;
  for (uint64_t p : paths) total += p;
  uint32_t max_npix = 0;
  for (uint32_t i = 0; i < n; ++i) max_npix = std::max(max_npix, tab[i].npix);
  std::vector<KernelCfg> group_cfg;        // knobs of each group
  std::vector<std::vector<uint32_t>> group_items;
  for (uint32_t i = 0; i < n; ++i) {
    const KernelCfg& c = items[i].scene->cfg;
    size_t g = 0;
    while (g < group_cfg.size() &&
           !(group_cfg[g].width == c.width && group_cfg[g].lds == c.lds && group_cfg[g].block == c.block))
      ++g;
    if (g == group_cfg.size()) {
      group_cfg.push_back(c);
      group_items.emplace_back();
    }
    group_items[g].push_back(i);
  }
  std::vector<ChunkDesc> chunks;
  std::vector<uint32_t> chunk_group;  // group of each chunk
  for (size_t g = 0; g < group_items.size(); ++g) {
    const std::vector<uint32_t>& gi = group_items[g];
    uint64_t gtotal = 0;
    for (uint32_t i : gi) gtotal += paths[i];
    const dgpu_scene* sg0 = items[gi[0]].scene;
    uint64_t C = std::min<uint64_t>(sg0->queue_tile ? sg0->queue_tile : ctx->batch_tile(2), 0x7FFFFFFFull);
    C = std::min<uint64_t>(C, 0xFFFFFFFFull - max_npix);  // q_begin + slots below 2^32
    if (gtotal > C) {
      uint64_t nch = (gtotal + C - 1) / C;
      C = (gtotal + nch - 1) / nch;
    }
    size_t k = 0;
    uint64_t local = 0;
    while (k < gi.size() && paths[gi[k]] == 0) ++k;
    while (k < gi.size()) {
      ChunkDesc cd;
      memset(&cd, 0, sizeof cd);
      while (k < gi.size() && cd.n_chunk < C && cd.n_segs < MAX_CHUNK_SEGS) {
        const uint32_t si = gi[k];
        uint64_t take = std::min<uint64_t>(paths[si] - local, C - cd.n_chunk);
        ChunkSeg& sg = cd.segs[cd.n_segs++];
        sg.scene = si;
        sg.slot_begin = cd.n_chunk;
        sg.pair_begin = cd.n_pairs;
        sg.local_begin = local;
        sg.local_end = local + take;
        sg.q_begin = (uint32_t)(local % tab[si].npix);
        sg.s_begin = tab[si].sample_begin + (uint32_t)(local / tab[si].npix);
        sg.n_pairs = (uint32_t)std::min<uint64_t>(tab[si].npix, take);
        cd.n_chunk += (uint32_t)take;
        cd.n_pairs += sg.n_pairs;
        local += take;
        if (local == paths[si]) {
          ++k;
          local = 0;
          while (k < gi.size() && paths[gi[k]] == 0) ++k;
        }
      }
      chunks.push_back(cd);
      chunk_group.push_back((uint32_t)g);
    }
  }
  if (chunks.empty()) return DGPU_OK;
  uint32_t cap = 0;
  for (const ChunkDesc& cd : chunks) cap = std::max(cap, cd.n_chunk);
  uint32_t any_rough = 0;
  for (uint32_t i = 0; i < n; ++i) any_rough |= items[i].scene->S.has_rough;
  const uint32_t shade_grid = ctx->shade_blocks(true, false, any_rough != 0);

  // per group: the extend and connect launch shapes; the spill area fits the largest
  std::vector<KernelCfg> group_ce(group_cfg), group_cc(group_cfg);
  size_t spill_threads = 0;
  for (size_t g = 0; g < group_cfg.size(); ++g) {
    group_ce[g].grid = ctx->grid_for(4, group_ce[g]);
    group_cc[g].grid = ctx->grid_for(5, group_cc[g]);
    spill_threads = std::max<size_t>(spill_threads,
                                     (size_t)std::max(group_ce[g].grid, group_cc[g].grid) * group_cfg[g].block);
  }
  WaveParams P;
  memset(&P, 0, sizeof P);
  P.S = s0->S;
  P.min_depth = o0.min_depth; P.max_depth = o0.max_depth; P.rr_depth = o0.rr_depth;
  P.only_direct = o0.only_direct; P.sample_all = 0; P.strategy = o0.direct_strategy;
  P.slots = 2;
  P.rough = any_rough;
  P.thr = (float4*)ctx->buf("thr", (size_t)cap * 16, &st);
  P.rad = (float4*)ctx->buf("rad", (size_t)cap * 16, &st);
  P.pend = (float4*)ctx->buf("pend", (size_t)cap * 2 * 16, &st);
  P.qo[0] = (float4*)ctx->buf("qo0", (size_t)cap * 16, &st);
  P.qo[1] = (float4*)ctx->buf("qo1", (size_t)cap * 16, &st);
  P.qd[0] = (float4*)ctx->buf("qd0", (size_t)cap * 16, &st);
  P.qd[1] = (float4*)ctx->buf("qd1", (size_t)cap * 16, &st);
  P.qpath[0] = (uint32_t*)ctx->buf("qp0", (size_t)cap * 4, &st);
  P.qpath[1] = (uint32_t*)ctx->buf("qp1", (size_t)cap * 4, &st);
  P.hit = (float4*)ctx->buf("hit", (size_t)cap * 16, &st);
  P.so = (float4*)ctx->buf("so", (size_t)cap * 2 * 16, &st);
  P.sd = (float4*)ctx->buf("sd", (size_t)cap * 2 * 16, &st);
  P.smeta = (uint2*)ctx->buf("smeta", (size_t)cap * 2 * 8, &st);
  P.counters = (uint32_t*)ctx->buf("counters", NUM_COUNTERS * 4, &st);
  P.stats = (unsigned long long*)ctx->buf("stats", STATS_WORDS * 8, &st);
  P.spill = (uint32_t*)ctx->buf("spill", spill_threads * stack_cap * 4, &st);
  P.stack_cap = stack_cap;
  // shadow-ray tests beside the next bounce's extension rays, as in dgpu_render
  WaveParams Pc;
  uint32_t* spill_c = (uint32_t*)ctx->buf("spill_c", spill_threads * stack_cap * 4, &st);
  hipStream_t stream_c = ctx->second_stream();
  if (!stream_c) return fail(DGPU_EDEVICE, "stream creation failed");
  BatchEntry* dtab = (BatchEntry*)ctx->buf("batch_tab", n * sizeof(BatchEntry), &st);
  ChunkDesc* dchunks = (ChunkDesc*)ctx->buf("batch_chunks", chunks.size() * sizeof(ChunkDesc), &st);
  uint32_t* dcur = (uint32_t*)ctx->buf("batch_cur", 64, &st);
  if (st != DGPU_OK) return st;
  P.batch = dtab;
  P.num_scenes = n;
  P.chunks = dchunks;
  P.chunk_cur = dcur;
  Pc = P;
  Pc.spill = spill_c;
  HIP_TRY(hipMemcpyAsync(dtab, tab.data(), n * sizeof(BatchEntry), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(dchunks, chunks.data(), chunks.size() * sizeof(ChunkDesc), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemsetAsync(dcur, 0, 4, stream));
  HIP_TRY(hipMemsetAsync(P.stats, 0, STATS_WORDS * 8, stream));
  // tab and chunks must stay alive until the copies above have run
  HIP_TRY(hipStreamSynchronize(stream));

  // chunk c: raygen -> (extend -> shade -> connect) x (max_depth + 1) -> finalize -> advance,
  // with its group's trace kernels; the kernels read the chunk index from the device.
  // Each bounce's connect runs on stream_c beside the next bounce's extend (the next
  // shade waits for it); every ordering point has an event of its own, so that a graph
  // captured from the two streams holds each dependency as recorded.
  size_t next_sync = 0;
  auto chunk_sequence = [&](size_t c) -> hipError_t {
    const KernelCfg& ce = group_ce[chunk_group[c]];
    const KernelCfg& cc = group_cc[chunk_group[c]];
    hipEvent_t ev_conn = nullptr;
    hipError_t e = launch_raygen(P, cap, true, stream);
    for (uint32_t b = 0; e == hipSuccess && b <= o0.max_depth; ++b) {
      e = launch_extend(P, ce, b, true, stream);
      if (e == hipSuccess && ev_conn) e = hipStreamWaitEvent(stream, ev_conn, 0);
      if (e == hipSuccess) e = launch_shade(P, cap, b, true, stream, shade_grid);
      hipEvent_t ev_shade = ctx->sync_event(next_sync++);
      ev_conn = ctx->sync_event(next_sync++);
      if (!ev_shade || !ev_conn) return hipErrorOutOfMemory;
      if (e == hipSuccess) e = hipEventRecord(ev_shade, stream);
      if (e == hipSuccess) e = hipStreamWaitEvent(stream_c, ev_shade, 0);
      if (e == hipSuccess) e = launch_connect(Pc, cc, b, true, stream_c);
      if (e == hipSuccess) e = hipEventRecord(ev_conn, stream_c);
    }
    if (e == hipSuccess && ev_conn) e = hipStreamWaitEvent(stream, ev_conn, 0);
    if (e == hipSuccess) e = launch_finalize(P, cap, true, stream);
    if (e == hipSuccess) e = launch_advance(dcur, stream);
    return e;
  };
  HIP_TRY(hipEventRecord(ctx->ev0, stream));
  if (use_graph) {
    // every chunk of every group in one graph: the whole batch is one graph launch
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    hipError_t ce_ = hipSuccess;
    for (size_t c = 0; c < chunks.size() && ce_ == hipSuccess; ++c) ce_ = chunk_sequence(c);
    hipError_t ee = hipStreamEndCapture(stream, &graph);
    if (ce_ != hipSuccess || ee != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return fail(DGPU_EDEVICE, "graph capture failed: %s", hipGetErrorString(ce_ != hipSuccess ? ce_ : ee));
    }
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return fail(DGPU_EDEVICE, "hipGraphInstantiate failed: %s", hipGetErrorString(ie));
    }
    hipError_t le = hipGraphLaunch(exec, stream);
    (void)hipStreamSynchronize(stream);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (le != hipSuccess) return fail(DGPU_EDEVICE, "hipGraphLaunch failed: %s", hipGetErrorString(le));
  } else {
    for (size_t c = 0; c < chunks.size(); ++c) HIP_TRY(chunk_sequence(c));
  }
  HIP_TRY(hipEventRecord(ctx->ev1, stream));
  unsigned long long hstats[STATS_WORDS] = {};
  HIP_TRY(hipMemcpyAsync(hstats, P.stats, sizeof hstats, hipMemcpyDeviceToHost, stream));
  std::vector<float> hfilm;
  if (!dev) {
    hfilm.resize(film_floats);
    HIP_TRY(hipMemcpyAsync(hfilm.data(), dfilm, film_floats * 4, hipMemcpyDeviceToHost, stream));
  }
  HIP_TRY(hipStreamSynchronize(stream));
  if (!dev) {
    size_t o = 0;
    for (uint32_t i = 0; i < n; ++i) {
      size_t np = tab[i].npix;
      for (size_t k = 0; k < 3 * np; ++k) items[i].film_rgb_sum[k] += hfilm[o + k];
      for (size_t k = 0; k < np; ++k) items[i].film_weight[k] += hfilm[o + 3 * np + k];
      o += 4 * np;
    }
  }
  if (out_stats) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    out_stats->paths = hstats[0];
    out_stats->extension_rays = hstats[1];
    out_stats->shadow_rays = hstats[2];
    out_stats->dropped_samples = hstats[3];
    for (int b = 0; b < 16; ++b) {
      out_stats->bounce_extension_rays[b] = hstats[4 + b];
      out_stats->bounce_shadow_rays[b] = hstats[20 + b];
    }
    out_stats->kernel_ms = ms;
  }
  (void)total;
  return DGPU_OK;
}

}  // extern "C"
