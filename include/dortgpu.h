/*
 * dortgpu.h -- C-ABI of the MI355X path-tracing core for dort.
 *
 * The library replaces the body of dort's CPU path tracer: `PathRenderer::render`
 * (src/dort/path_renderer.cpp:11-21) and everything it calls per sample
 * (scene intersection, BSDFs, lights, film accumulation).  The seam in dort is
 * the abstract `Renderer` (include/dort/renderer.hpp:12-34), selected by
 * `lua_render_make` (src/dort/lua_render.cpp:102-221).  The host adapter
 * `dort_amd::GpuPathRenderer` (dort_amd/host/gpu_path_renderer.hpp) is such a
 * Renderer: it takes the scene as a typed `dgpu_scene_desc` recorded next to the
 * dort::Scene and calls the functions below.
 *
 * Conventions
 *  - C99, no exceptions cross this boundary.  Every call returns a dgpu_status;
 *    `dgpu_last_error()` gives a thread-local message for the last failure.
 *  - Handles are opaque and library-owned.  Descriptor arrays are host memory
 *    and are copied; the caller keeps ownership.
 *  - Ray/hit/film buffers are host or device pointers, chosen per call by
 *    DGPU_BUF_DEVICE in `flags`.  Device pointers must be allocations on the
 *    context's device (e.g. torch tensors on that device).
 *  - Up to three HIP streams per context, all of default priority (renders test
 *    shadow rays on the second beside the first's extension rays; the light tracer
 *    tests alternate stages' camera connections on the second and third, BDPT runs
 *    its strategy kernels on all three); calls on one context are serialised by an
 *    internal mutex.  Separate contexts are
 *    independent.
 *  - Rendering is deterministic for a given (scene, opts.seed, sample range):
 *    random numbers are drawn from a counter-based generator keyed by
 *    (seed, pixel, sample, bounce, dimension), never by launch geometry.
 */
#ifndef DORTGPU_H
#define DORTGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGPU_ABI_VERSION 4u

typedef enum {
  DGPU_OK = 0,
  DGPU_EINVAL = 1,       /* bad argument (null handle, zero size, bad index ...) */
  DGPU_ENOMEM = 2,       /* device or host allocation failed */
  DGPU_EDEVICE = 3,      /* HIP runtime error (no device, launch failure ...) */
  DGPU_EUNSUPPORTED = 4, /* scene feature not supported by the GPU path */
  DGPU_ECANCELLED = 5    /* cancel flag observed between bounces */
} dgpu_status;

typedef struct dgpu_ctx dgpu_ctx;
typedef struct dgpu_scene dgpu_scene;

/* ---- scene description ------------------------------------------------- */

enum {
  DGPU_MAT_NONE = 0,          /* no BSDF (e.g. the black Lambert dort puts on area lights) */
  DGPU_MAT_LAMBERT = 1,       /* lambert_material.cpp:5-35, two-sided */
  DGPU_MAT_GGX_CONDUCTOR = 2, /* GGX microfacet conductor (no dort counterpart) */
  DGPU_MAT_DIELECTRIC = 3,    /* dielectric_material.cpp:32-104, delta reflect/refract */
  DGPU_MAT_ROUGH_DIELECTRIC = 4 /* dort's RoughDielectricBxdf as dort runs it
                                   (rough_dielectric_material.cpp:7-185): GGX of alpha =
                                   roughness; sampling reflects with the Fresnel probability
                                   and a transmission pick yields nothing; eval gives the
                                   reflection or transmission term by hemisphere; its pdf is
                                   0 (so MIS weights are dort's); roughness < 0.001 is the
                                   smooth dielectric.  ior_*, reflect_tint, transmit_tint as
                                   for the dielectric */
};

typedef struct {
  uint32_t type;           /* DGPU_MAT_* */
  int32_t albedo_tex;      /* Lambert: texture index (uv lookup) or -1 for constant albedo */
  float albedo[3];         /* Lambert albedo / conductor tint */
  float roughness;         /* GGX alpha (no remapping, microfacet.cpp:10-12): conductor, rough dielectric */
  float eta[3];            /* conductor complex IOR, real part (per RGB channel) */
  float k[3];              /* conductor complex IOR, imaginary part */
  float ior_inside;        /* dielectric, rough dielectric */
  float ior_outside;       /* (dort's Lua default is 1, lua_material.cpp:118) */
  float reflect_tint[3];
  float transmit_tint[3];
} dgpu_material;

typedef struct {
  uint32_t triangle;       /* index into the scene triangles (the emitting shape) */
  float radiance[3];       /* one-sided emitted radiance, diffuse_light.cpp:83-93 */
} dgpu_area_light;

typedef struct {
  uint32_t width, height;  /* lat-long map, rows = theta from `up` (environment_light.cpp:163-186) */
  const float* rgb;        /* width*height*3, row-major */
  float up[3];
  float fwd[3];
  float scale[3];
} dgpu_env_light;

/* How a texture is addressed at a surface point (texture_maps.cpp:4-15). */
enum {
  DGPU_TEXMAP_UV = 0,  /* uv_texture_map: the surface's uv */
  DGPU_TEXMAP_XY = 1   /* xy_texture_map(tex_to_world): x, y of world_to_texture applied to the point */
};

/* An image texture composed with its mapping: ImageTexture (image_texture.cpp:5-11), nearest
 * texel, coordinates wrapped to [0, 1), row y = floor(height * v) from the image's first row. */
typedef struct {
  uint32_t width, height;
  const float* rgb;        /* width*height*3 fp32 texels, or NULL when rgb8 is given */
  const uint8_t* rgb8;     /* width*height*3 8-bit texels (dort's Image<PixelRgb8>): k reads as
                              k / 256 (PixelRgb8::to_rgb, image.hpp:66-71); used when rgb is NULL */
  uint32_t mapping;        /* DGPU_TEXMAP_* */
  float world_to_texture[12]; /* DGPU_TEXMAP_XY: row-major 3x4, the inverse of the map's
                                 tex_to_world (Transform::apply_inv of texture_maps.cpp:12) */
} dgpu_texture;

typedef struct {
  float camera_to_world[12]; /* row-major 3x4 affine (dort Transform, pinhole_camera.cpp:6-11) */
  float fov;                 /* spans the longer film dimension (pinhole_camera.cpp:163-181) */
} dgpu_camera;

typedef struct {
  uint32_t num_vertices;
  uint32_t num_triangles;
  const float* positions;      /* num_vertices*3 */
  const uint32_t* indices;     /* num_triangles*3 */
  const float* uvs;            /* num_vertices*2, or NULL: (0,0),(1,0),(0,1) per triangle (triangle.cpp:89-93) */
  const uint32_t* tri_material;/* num_triangles material indices */
  uint32_t num_materials;
  const dgpu_material* materials;
  uint32_t num_area_lights;
  const dgpu_area_light* area_lights;
  const dgpu_env_light* env;   /* NULL if none */
  const float* infinite_radiance; /* NULL, or rgb of an InfiniteLight (infinite_light.cpp) */
  uint32_t num_textures;
  const dgpu_texture* textures;
  dgpu_camera camera;
} dgpu_scene_desc;

/* BVH / kernel knobs (the autotuner's search space). Zero fields take defaults. */
typedef struct {
  uint32_t width;          /* 2, 4 or 8 (default 4) */
  uint32_t max_leaf_size;  /* triangles per leaf, 1..8 (default 4) */
  uint32_t sah_bins;       /* binned-SAH bins (default 16) */
  uint32_t stack_lds;      /* traversal stack entries kept in LDS per lane (default 16) */
  uint32_t block_size;     /* threads per traversal workgroup: 64,128,256 (default 128) */
  uint32_t queue_tile;     /* paths per wavefront chunk (default: sized from free memory) */
  float sah_node_cost;     /* SAH cost of one node visit in triangle tests (default 1) */
} dgpu_bvh_opts;

/* ---- rays and hits (SoA, trace_* calls) -------------------------------- */

typedef struct {
  const float* ox; const float* oy; const float* oz;
  const float* dx; const float* dy; const float* dz;
  const float* tmin; const float* tmax;
} dgpu_rays;

typedef struct {
  float* t;          /* hit distance, INFINITY when missed */
  int32_t* prim;     /* scene triangle index, -1 when missed */
  float* b1;         /* barycentric weight of vertex 1 */
  float* b2;         /* barycentric weight of vertex 2 */
} dgpu_hits;

enum {
  DGPU_BUF_DEVICE = 1u,     /* buffers are device pointers */
  DGPU_RENDER_PROFILE = 2u, /* dgpu_render: time every kernel launch with HIP events */
  DGPU_BATCH_NO_GRAPH = 4u, /* dgpu_render_batch: launch kernels directly, no hipGraph */
  DGPU_TRACE_BOUNDED = 8u,  /* dgpu_trace_*: every ray origin lies within the scene's range (no
                               coordinate beyond max |vertex or camera coordinate|), which
                               allows the path tracer's faster box test */
  DGPU_ADJOINT_MEGAKERNEL = 16u, /* dgpu_adjoint: one thread per path traces and replays it in
                                    one kernel, instead of the two wavefront passes */
  DGPU_BDPT_MEGAKERNEL = 32u     /* dgpu_render_bdpt: one thread per path sample walks both
                                    subpaths and forms every strategy, instead of the
                                    wavefront stages */
};

/* ---- rendering ----------------------------------------------------------- */

enum { DGPU_DIRECT_MIS = 0, DGPU_DIRECT_BSDF = 1, DGPU_DIRECT_LIGHT = 2 };

typedef struct {
  uint32_t x_res, y_res;
  uint32_t sample_begin;   /* iterations [sample_begin, sample_end); iteration 0 samples pixel centres */
  uint32_t sample_end;
  uint64_t seed;
  uint32_t min_depth;      /* path_renderer.hpp:16-21, lua_render.cpp:125-145 */
  uint32_t max_depth;
  uint32_t rr_depth;       /* RR when bounces >= rr_depth && bounces > min_depth; dort's rule is 2 */
  uint32_t only_direct;
  uint32_t sample_all_lights;
  uint32_t direct_strategy;/* DGPU_DIRECT_* */
  uint32_t flags;          /* DGPU_BUF_DEVICE: film pointers are device memory */
  uint32_t scene_id;       /* mixed into the RNG key; batch renders use the scene's position */
} dgpu_render_opts;

typedef struct {
  uint64_t paths;           /* paths traced */
  uint64_t extension_rays;  /* closest-hit rays */
  uint64_t shadow_rays;     /* any-hit rays (NEE light + NEE BSDF strategies) */
  uint64_t dropped_samples; /* non-finite or negative samples dropped (renderer.cpp:36-41) */
  double kernel_ms;         /* GPU time of the render, HIP events */
  /* With DGPU_RENDER_PROFILE: GPU time and launch count per kernel kind. */
  double extend_ms, shade_ms, connect_ms, other_ms;
  uint32_t extend_launches, shade_launches, connect_launches, other_launches;
  /* Queue lengths per bounce (bounces past 15 are added to [15]). */
  uint64_t bounce_extension_rays[16];
  uint64_t bounce_shadow_rays[16];
} dgpu_stats;

/* ---- context ------------------------------------------------------------- */

uint32_t dgpu_abi_version(void);
const char* dgpu_last_error(void);

/* Creates a context on HIP device `device` with its own stream. */
dgpu_status dgpu_create(int device, dgpu_ctx** out);
void dgpu_destroy(dgpu_ctx* ctx);

/* Device memory on the context's device, for callers that keep rays, hits or films
 * resident in HBM (DGPU_BUF_DEVICE).  kind: 0 host->device, 1 device->host,
 * 2 device->device.  Copies are synchronous. */
dgpu_status dgpu_device_alloc(dgpu_ctx* ctx, size_t bytes, void** out);
dgpu_status dgpu_device_free(dgpu_ctx* ctx, void* ptr);
dgpu_status dgpu_copy(dgpu_ctx* ctx, void* dst, const void* src, size_t bytes, int kind);

/* ---- scenes -------------------------------------------------------------- */

/* Builds the BVH on the host (binned SAH, flattened to the GPU node layout) and
 * uploads geometry, materials, lights, textures and the environment map. */
dgpu_status dgpu_scene_create(dgpu_ctx* ctx, const dgpu_scene_desc* desc,
                              const dgpu_bvh_opts* opts, dgpu_scene** out);
void dgpu_scene_destroy(dgpu_scene* scene);

typedef struct {
  uint32_t nodes;          /* BVH nodes */
  uint32_t leaves;
  uint32_t depth;
  uint32_t width;
  uint64_t node_bytes;
  uint64_t triangle_bytes;
  double build_ms;         /* host BVH build + flatten */
  double upload_ms;
} dgpu_scene_info;

dgpu_status dgpu_scene_get_info(const dgpu_scene* scene, dgpu_scene_info* out);

/* ---- intersection (Scene::intersect / Scene::intersect_p, scene.cpp:8-18) - */

/* Closest hit per ray; watertight FP32 triangle test, two-sided.  flags: DGPU_BUF_DEVICE,
 * DGPU_TRACE_BOUNDED (same results, faster; only for origins within the scene's range). */
dgpu_status dgpu_trace_closest(dgpu_scene* scene, const dgpu_rays* rays, uint32_t n,
                               dgpu_hits* out, uint32_t flags, double* kernel_ms);
/* Any hit per ray in [tmin, tmax]; occluded[i] = 1 if some triangle is hit. */
dgpu_status dgpu_trace_any(dgpu_scene* scene, const dgpu_rays* rays, uint32_t n,
                           uint32_t* occluded, uint32_t flags, double* kernel_ms);

/* dgpu_trace_closest that allocates no memory.  No counterpart in dort: it is for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Rays and hits are device buffers (DGPU_BUF_DEVICE must be set;
 * DGPU_TRACE_BOUNDED may be); the hits equal dgpu_trace_closest's.  The kernel's scratch
 * (its work counter and the traversal stack's spill slices) is a fixed 32 MiB array on
 * each GPU, held from the library's first use until the process ends; the persistent grid
 * shrinks to fit it, and a scene whose stack would not fit one workgroup is refused with
 * DGPU_EUNSUPPORTED.  Every call on a GPU shares that GPU's array, so calls that overlap on
 * one GPU would race: the library serialises them per GPU and returns once the kernel is
 * done. */
dgpu_status dgpu_trace_closest_noalloc(dgpu_scene* scene, const dgpu_rays* rays, uint32_t n,
                                       dgpu_hits* out, uint32_t flags, double* kernel_ms);

/* Replace the texels of texture `tex` (width*height*3 fp32, host memory) or one
 * material, without rebuilding the BVH (inverse rendering, finite differences). */
dgpu_status dgpu_scene_update_texture(dgpu_scene* scene, uint32_t tex, const float* rgb);
dgpu_status dgpu_scene_update_material(dgpu_scene* scene, uint32_t material, const dgpu_material* m);

/* ---- path tracing (PathRenderer::render, path_renderer.cpp:11-206) -------- */

/* Adds the samples of iterations [sample_begin, sample_end) to the film:
 * film_rgb_sum[3*(y*x_res+x)+c] += L, film_weight[y*x_res+x] += 1 per sample
 * (box filter of radius 0.5, film.cpp:16-30).  `cancel` (host memory, may be
 * NULL) is polled between bounces. */
dgpu_status dgpu_render(dgpu_scene* scene, const dgpu_render_opts* opts,
                        float* film_rgb_sum, float* film_weight,
                        dgpu_stats* out_stats, const volatile int* cancel);

/* dort's light tracer, LightRenderer::render (src/dort/light_renderer.cpp:13-181, the
 * "lt" renderer of lua_render.cpp:146-150): per iteration x_res*y_res paths start on a
 * light picked by power; every vertex is connected to the camera and splatted into
 * the film with the 1 px box filter.  Uses x_res, y_res, sample_begin/sample_end
 * (iterations), seed, scene_id, flags and min_depth/max_depth (paths of min_depth + 2
 * .. max_depth + 2 vertices, as dort maps them); Russian roulette is dort's (from the
 * third bounce, survival clamped to [0.1, 0.9]).  film_splat_rgb (x_res*y_res*3, host
 * or device memory with DGPU_BUF_DEVICE) receives += the splatted contributions;
 * dort's image is the sum times 1 / (total iterations) (Film::to_image's splat_scale).
 * Splats are float atomics, so sums vary in the last bits from run to run. */
dgpu_status dgpu_render_light(dgpu_scene* scene, const dgpu_render_opts* opts, float* film_splat_rgb,
                              dgpu_stats* out_stats, const volatile int* cancel);

/* dort's bidirectional path tracer, BdptRenderer::render (src/dort/bdpt_renderer.cpp:9-749,
 * the "bdpt" renderer of lua_render.cpp:151-160): per iteration and pixel one light
 * subpath (a light picked by power) and one camera subpath through pixel + a uniform
 * jitter, each of at most max_depth + 2 vertices with dort's Russian roulette (from the
 * third vertex, survival min(mean throughput scale, 0.9)); every strategy of s light and
 * t camera vertices with min_depth + 2 <= s + t <= max_depth + 2 is weighted with dort's
 * balance heuristic.  Strategies with t >= 2 form the pixel's sample, added like
 * dgpu_render's (film_rgb_sum += L, film_weight += 1; samples exactly on a pixel edge,
 * and non-finite ones, are dropped); strategies with t = 1 (light subpath to the camera,
 * only with use_t1_paths) are splatted into film_splat_rgb with the 1 px box filter.
 * dort's image is film_rgb_sum / film_weight + film_splat_rgb / (total iterations).  The
 * strategy s = t = 1 has weight 0 for area and background lights and is not formed.
 * Uses x_res, y_res, sample_begin/sample_end, seed, scene_id, flags (DGPU_BUF_DEVICE:
 * all three films are device memory; DGPU_BDPT_MEGAKERNEL: one thread per path sample
 * instead of the wavefront stages, the same subpaths and strategies), min_depth and
 * max_depth (< 30); the other options are the path tracer's.  A scene without lights
 * adds nothing (bdpt_renderer.cpp:11).  The wavefront stages keep per path of a chunk
 * 2 (max_depth + 2) 64 B of subpath vertices and 264 B of walk state, and per path of the
 * part of the chunk whose strategies are formed at once 64 B per strategy slot
 * ((max_depth + 2) (max_depth + 3) / 2 slots); both are sized from free device memory,
 * at most 32 GiB together.  Splats are float atomics, so their sums vary in the last
 * bits from run to run. */
dgpu_status dgpu_render_bdpt(dgpu_scene* scene, const dgpu_render_opts* opts, uint32_t use_t1_paths,
                             float* film_rgb_sum, float* film_weight, float* film_splat_rgb,
                             dgpu_stats* out_stats, const volatile int* cancel);

/* ---- batched rendering (no dort counterpart) -------------------------------
 * Renders many scenes in one wavefront: every scene's paths go into shared queues, and
 * the bounce loop of every chunk (raygen, extend/shade/connect per bounce, film) is
 * captured in one hipGraph, launched once; the chunk index advances on the device.
 * Scenes may differ in triangles, resolution, sample range, seed, BVH width and kernel
 * knobs (dgpu_scene_create's width, stack_lds and block_size): items whose scenes share
 * the knobs form a group, and each group's paths are cut into chunks of their own,
 * traced by that group's kernel variant.  Items must share the depth / strategy options
 * and a context.  Each item's film is accumulated like dgpu_render's, so batch results
 * equal per-scene renders.  n is in [1, 255].  A group's chunks hold at most its first
 * scene's queue_tile paths, or by default equal shares of the group of at most 32 GiB of
 * path state each (about 148M paths); a chunk spans at most 64 scenes.
 * flags: DGPU_BUF_DEVICE (films are device pointers), DGPU_BATCH_NO_GRAPH.  Device films
 * are added to without atomics, so no two of them (nor an item's colour and weight
 * arrays) may overlap: such a batch fails with DGPU_EINVAL.  Host films may be shared
 * (each item is staged separately and added at the end). */
typedef struct {
  dgpu_scene* scene;
  dgpu_render_opts opts;
  float* film_rgb_sum;
  float* film_weight;
} dgpu_batch_item;

dgpu_status dgpu_render_batch(dgpu_ctx* ctx, const dgpu_batch_item* items, uint32_t n,
                              uint32_t flags, dgpu_stats* out_stats);

/* ---- adjoint path replay (no dort counterpart) ----------------------------
 * Gradient of L = sum_px dot(dL_dimage[px], I[px]), where I is the film mean of
 * iterations [sample_begin, sample_end) rendered with `opts`, with respect to the
 * texels of texture 0 (d_albedo_texels: width*height*3, may be NULL) and every
 * material's GGX roughness (d_roughness: num_materials, may be NULL).  Each path is
 * traced twice from its counter-based random numbers (forward, then replay with
 * gradient deposits): by default as two wavefront passes over a chunk of paths (the
 * first keeps each path's radiance, 16 B per path), with DGPU_ADJOINT_MEGAKERNEL in
 * opts->flags by one thread per path; nothing is taped.  Outputs are accumulated (+=).
 * With DGPU_BUF_DEVICE in opts->flags all three buffers are device pointers.
 * image_out (optional, same memory kind) receives += the per-pixel mean radiance of
 * the replayed paths. */
dgpu_status dgpu_adjoint(dgpu_scene* scene, const dgpu_render_opts* opts, const float* dL_dimage,
                         float* d_albedo_texels, float* d_roughness, float* image_out,
                         dgpu_stats* out_stats);

/* ---- scene and image files (host; dort_amd/csrc/io.cpp) -------------------- */

/* PLY mesh (ascii, binary little or big endian), as dort's read_ply_callbacks /
 * read_ply_to_mesh read it (src/dort/ply_mesh.cpp:9-103): vertex x, y, z and each face's
 * vertex_indices list as the fan (i0, i1, i2), (i0, i2, i3), ...  The arrays are
 * allocated by the library (release them with dgpu_free).  Unlike dort, an index out
 * of range is an error (DGPU_EINVAL) instead of undefined behaviour. */
dgpu_status dgpu_read_ply(const char* path, float** positions, uint32_t* num_vertices,
                          uint32_t** indices, uint32_t* num_triangles);
void dgpu_free(void* p);

/* Radiance RGBE images of 3-channel fp32 pixels, row-major from the top, as dort's
 * read_image_f / write_image_rgbe (src/dort/image.cpp:28-67, stb_image) read and write
 * them: "-Y H +X W" scanlines, flat or run-length encoded. */
dgpu_status dgpu_read_rgbe(const char* path, float** rgb, uint32_t* width, uint32_t* height);
dgpu_status dgpu_write_rgbe(const char* path, const float* rgb, uint32_t width, uint32_t height);

/* PNG images as dort's read_image_8 / write_image_png (src/dort/image.cpp:14-26, 42-51,
 * stb_image) read and write them: any colour type and bit depth (and Adam7) read into
 * 3-byte RGB pixels, row-major from the top, converted as stb converts (16-bit: high
 * byte; grey of 1/2/4 bits scaled to 0..255; palette looked up; grey replicated; alpha
 * dropped); 3-byte RGB written as an 8-bit colour-type-2 PNG.  The pixels read are
 * allocated by the library (release them with dgpu_free). */
dgpu_status dgpu_read_png(const char* path, uint8_t** rgb, uint32_t* width, uint32_t* height);
dgpu_status dgpu_write_png(const char* path, const uint8_t* rgb, uint32_t width, uint32_t height);

/* ---- image comparison (GPU; dort_amd/csrc/compare.hip) --------------------- */

/* dort's test_convergence (src/dort/convergence_test.cpp:10-101) over width*height*3 fp32
 * images (host memory, or device memory with DGPU_BUF_DEVICE): *passed = 1 when every
 * tile of every depth passes, else 0 with the first failing tile (smallest depth, then
 * row-major) described in msg as dort describes it. */
dgpu_status dgpu_test_convergence(dgpu_ctx* ctx, const float* reference, const float* tested,
                                  uint32_t width, uint32_t height, int32_t min_tile_size,
                                  float variation, float bias, float p_value, uint32_t flags,
                                  int32_t* passed, char* msg, size_t msg_size);
/* dort's bias_variance (src/dort/lua_image.cpp:179-211) over [x0, x1) x [y0, y1): the sum
 * of (reference - tested) and the sum of its squares over the pixels' channels. */
dgpu_status dgpu_bias_variance(dgpu_ctx* ctx, const float* reference, const float* tested,
                               uint32_t width, uint32_t height, uint32_t x0, uint32_t y0,
                               uint32_t x1, uint32_t y1, uint32_t flags, double* bias,
                               double* variance);

#ifdef __cplusplus
}
#endif

#endif /* DORTGPU_H */
