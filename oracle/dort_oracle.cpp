// dort_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// The oracle: a CPU restatement of dort's plain path tracer (`renderer = "pt"`)
// and of everything it calls per sample, written fresh from the dort sources
// cited in orc_core.inc.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library, and only as the checker / CPU baseline;
// the product (dort_amd/libdortgpu.so) never links or calls it.
//
// Exposed over a small C ABI for ctypes:
//   orc_scene_create/destroy   scene from the same dgpu_scene_desc as the GPU path
//   orc_trace                  closest / any hit through dort's BVH, or brute force
//   orc_render                 PathRenderer::render, tiles x iterations on a thread pool
//   orc_adjoint64              FP64 path-replay adjoint (albedo texels, GGX roughness)
//   orc_bsdf_*, orc_env_*      per-function probes for closed-form tests
//   orc_fresnel_*, orc_philox  the Fresnel terms and the counter-based generator alone
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../include/dgpu_rng.h"
#include "../include/dortgpu.h"

namespace orc_util {
// ThreadPool::loop / parallel_for (thread_pool.cpp:49-73): a shared job counter.
void parallel_jobs(uint32_t threads, uint64_t jobs, const std::function<void(uint64_t, uint32_t)>& fn) {
  if (threads <= 1 || jobs <= 1) {
    for (uint64_t j = 0; j < jobs; ++j) fn(j, 0);
    return;
  }
  std::atomic<uint64_t> next{0};
  std::vector<std::thread> ts;
  for (uint32_t t = 0; t < threads; ++t) {
    ts.emplace_back([&, t]() {
      for (;;) {
        uint64_t j = next.fetch_add(1);
        if (j >= jobs) break;
        fn(j, t);
      }
    });
  }
  for (auto& t : ts) t.join();
}

uint32_t round_up_pow2(uint32_t x) {
  x = x - 1;
  x |= x >> 1; x |= x >> 2; x |= x >> 4; x |= x >> 8; x |= x >> 16;
  return x + 1;
}

// Renderer::layout_tiles (renderer.cpp:9-24)
void layout_tiles(uint32_t threads, uint32_t xr, uint32_t yr, uint32_t& tx, uint32_t& ty) {
  uint32_t approx = std::max(1u, std::max(threads * 8u, xr * yr / (32u * 1024u)));
  tx = round_up_pow2(approx);
  ty = 1;
  while (xr * ty < yr * tx) { tx /= 2; ty *= 2; }
}

}  // namespace orc_util

namespace orc_f {
using real = float;
#include "orc_core.inc"
}  // namespace orc_f

namespace orc_d {
using real = double;
#include "orc_core.inc"
}  // namespace orc_d

namespace {

struct OrcScene {
  orc_f::Scene* f = nullptr;
  orc_d::Scene* d = nullptr;
};

}  // namespace

extern "C" {

void* orc_scene_create(const dgpu_scene_desc* d, uint32_t split_method, uint32_t leaf_size,
                       uint32_t max_leaf_size, uint32_t fp64) {
  if (!d) return nullptr;
  OrcScene* s = new OrcScene();
  s->f = orc_f::load_scene(d, (int)split_method, leaf_size, max_leaf_size);
  if (fp64) s->d = orc_d::load_scene(d, (int)split_method, leaf_size, max_leaf_size);
  return s;
}

void orc_scene_destroy(void* p) {
  OrcScene* s = (OrcScene*)p;
  if (!s) return;
  delete s->f;
  delete s->d;
  delete s;
}

void orc_scene_info(void* p, uint32_t* nodes, uint32_t* depth, uint32_t* lights) {
  OrcScene* s = (OrcScene*)p;
  *nodes = (uint32_t)s->f->bvh.nodes.size();
  *depth = s->f->bvh.depth;
  *lights = (uint32_t)s->f->lights.size();
}

// rays: 8 pointers (ox, oy, oz, dx, dy, dz, tmin, tmax).  any != 0: occluded[i].
void orc_trace(void* p, uint32_t n, const float* const* rays, int any, int brute, uint32_t threads,
               float* t, int32_t* prim, float* b1, float* b2, uint32_t* occluded, uint64_t* visited) {
  using namespace orc_f;
  const Scene& sc = *((OrcScene*)p)->f;
  std::atomic<uint64_t> vis{0};
  const uint64_t chunk = 4096;
  orc_util::parallel_jobs(threads, (n + chunk - 1) / chunk, [&](uint64_t job, uint32_t) {
    uint64_t lv = 0;
    for (uint64_t i = job * chunk; i < std::min<uint64_t>(n, (job + 1) * chunk); ++i) {
      Ray r;
      r.o = V3(rays[0][i], rays[1][i], rays[2][i]);
      r.d = V3(rays[3][i], rays[4][i], rays[5][i]);
      r.tmin = rays[6][i]; r.tmax = rays[7][i];
      if (any) {
        bool hit = false;
        if (brute) {
          for (uint32_t k = 0; k < sc.mesh.ntri && !hit; ++k) {
            real th, u, v;
            hit = tri_hit(sc.mesh, k, r, th, u, v);
          }
        } else {
          hit = scene_intersect_p(sc, r, &lv);
        }
        occluded[i] = hit ? 1u : 0u;
      } else {
        Hit h;
        bool hit;
        if (brute) {
          hit = false;
          for (uint32_t k = 0; k < sc.mesh.ntri; ++k) {
            real th, u, v;
            if (tri_hit(sc.mesh, k, r, th, u, v)) {
              hit = true; r.tmax = th; h.t = th; h.prim = (int32_t)k; h.b1 = u; h.b2 = v;
            }
          }
        } else {
          hit = scene_intersect(sc, r, h, &lv);
        }
        t[i] = hit ? h.t : INFINITY;
        prim[i] = hit ? h.prim : -1;
        b1[i] = hit ? h.b1 : 0.f;
        b2[i] = hit ? h.b2 : 0.f;
      }
    }
    vis += lv;
  });
  if (visited) *visited = vis.load();
}

// rng_mode 0: dort's mt19937 streams; 1: the counter-based generator shared with the GPU path.
// counters: paths, closest-hit rays, any-hit rays, dropped samples.
int orc_render(void* p, const dgpu_render_opts* o, int rng_mode, uint32_t threads, float* rgb_sum,
               float* weight, uint64_t* counters) {
  OrcScene* s = (OrcScene*)p;
  if (!s || !o || o->sample_end < o->sample_begin) return 1;
  return orc_f::render(*s->f, o, rng_mode, threads, rgb_sum, weight, counters);
}


// FP64 path-replay adjoint (counter-based RNG): gradient of sum_px dldi . I(px) with
// respect to the texels of texture 0 (RGB) and every material's GGX roughness.
// image (optional) receives the FP64 forward image (mean over samples).
int orc_adjoint64(void* p, const dgpu_render_opts* o, uint32_t threads, const float* dldi,
                  double* g_tex, uint32_t n_texels, double* g_rough, uint32_t n_mats, float* image) {
  OrcScene* s = (OrcScene*)p;
  if (!s || !s->d || !o) return 1;
  return orc_d::adjoint(*s->d, o, threads, dldi, g_tex, n_texels, g_rough, n_mats, image);
}

// FP32 variant of the same replay (dort arithmetic), for like-for-like GPU comparisons.
int orc_adjoint32(void* p, const dgpu_render_opts* o, uint32_t threads, const float* dldi,
                  double* g_tex, uint32_t n_texels, double* g_rough, uint32_t n_mats, float* image) {
  OrcScene* s = (OrcScene*)p;
  if (!s || !o) return 1;
  return orc_f::adjoint(*s->f, o, threads, dldi, g_tex, n_texels, g_rough, n_mats, image);
}

// Bxdf probes in the local shading frame (n = +z); albedo is the material's constant.
void orc_bsdf_sample(const dgpu_material* m, uint32_t n, const float* wo, const float* uvc,
                     uint32_t request, float* wi, float* f, float* pdf, uint32_t* flags) {
  using namespace orc_f;
  Bsdf b = probe_bsdf(m);
  for (uint32_t i = 0; i < n; ++i) {
    V3 w(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), out(0, 0, 0);
    real pd; uint32_t fl;
    Spec fv = b.sample_light_f(w, request, out, pd, fl, uvc[3 * i], uvc[3 * i + 1], uvc[3 * i + 2]);
    wi[3 * i] = out.x; wi[3 * i + 1] = out.y; wi[3 * i + 2] = out.z;
    f[3 * i] = fv.r; f[3 * i + 1] = fv.g; f[3 * i + 2] = fv.b;
    pdf[i] = pd; flags[i] = fl;
  }
}

void orc_bsdf_eval(const dgpu_material* m, uint32_t n, const float* wi, const float* wo,
                   uint32_t request, float* f, float* pdf) {
  using namespace orc_f;
  Bsdf b = probe_bsdf(m);
  for (uint32_t i = 0; i < n; ++i) {
    V3 a(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), c(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]);
    Spec fv = b.eval_f(a, c, request);
    f[3 * i] = fv.r; f[3 * i + 1] = fv.g; f[3 * i + 2] = fv.b;
    pdf[i] = b.light_f_pdf(a, c, request);
  }
}

// Environment-light probes: sample_dir from (u, v), and pdf / radiance of directions.
void orc_env_sample(void* p, uint32_t n, const float* uv, float* dir, float* pdf, float* rad) {
  using namespace orc_f;
  const Scene& sc = *((OrcScene*)p)->f;
  for (uint32_t i = 0; i < n; ++i) {
    real eu, ev, pd;
    V3 d = sc.env.sample_dir(uv[2 * i], uv[2 * i + 1], eu, ev, pd);
    Spec r = sc.env.radiance(eu, ev);
    dir[3 * i] = d.x; dir[3 * i + 1] = d.y; dir[3 * i + 2] = d.z;
    pdf[i] = pd;
    rad[3 * i] = r.r; rad[3 * i + 1] = r.g; rad[3 * i + 2] = r.b;
  }
}

void orc_env_eval(void* p, uint32_t n, const float* dir, float* pdf, float* rad) {
  using namespace orc_f;
  const Scene& sc = *((OrcScene*)p)->f;
  for (uint32_t i = 0; i < n; ++i) {
    V3 d(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
    pdf[i] = sc.env.dir_pdf(d);
    Spec r = sc.env.background(d);
    rad[3 * i] = r.r; rad[3 * i + 1] = r.g; rad[3 * i + 2] = r.b;
  }
}

// Fresnel probes for closed-form tests.  Conductor: unpolarised reflectance for complex
// IOR eta + i k at cos_i, in the renderer's float arithmetic or in double.  Dielectric:
// dort's reflectance of a ray going from ior_refl into ior_trans (fresnel_dielectric with
// the refracted cosine of refract_dielectric, fresnel.cpp:5-35; 1 under total internal
// reflection), as the dielectric lobe computes it.
void orc_fresnel_conductor(uint32_t n, const double* cos_i, double eta, double k, int fp64, double* out) {
  for (uint32_t i = 0; i < n; ++i)
    out[i] = fp64 ? orc_d::fresnel_conductor1(cos_i[i], eta, k)
                  : (double)orc_f::fresnel_conductor1((float)cos_i[i], (float)eta, (float)k);
}

void orc_fresnel_dielectric(uint32_t n, const double* cos_i, double ior_refl, double ior_trans, int fp64,
                            double* out) {
  for (uint32_t i = 0; i < n; ++i) {
    if (fp64) {
      using namespace orc_d;
      real c = cos_i[i];
      V3 w(std::sqrt(std::max(0.0, 1.0 - c * c)), 0.0, c), t;
      out[i] = refract_dielectric(w, t, ior_refl, ior_trans) ? fresnel_dielectric(c, std::abs(t.z), ior_refl, ior_trans)
                                                            : 1.0;
    } else {
      using namespace orc_f;
      real c = (real)cos_i[i];
      V3 w(std::sqrt(std::max(0.f, 1.f - c * c)), 0.f, c), t;
      out[i] = refract_dielectric(w, t, (real)ior_refl, (real)ior_trans)
                   ? fresnel_dielectric(c, std::abs(t.z), (real)ior_refl, (real)ior_trans)
                   : 1.0;
    }
  }
}

// Philox4x32-10 of include/dgpu_rng.h (the generator the GPU path and the oracle's counter
// mode draw from) on one counter and key, for known-answer tests.
void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  dgpu_u32x4 r = dgpu_philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

// Light table: kinds (0 area, 1 env, 2 infinite) and the pick CDF (n+1 entries).
uint32_t orc_lights(void* p, int32_t* kinds, float* cdf) {
  const orc_f::Scene& sc = *((OrcScene*)p)->f;
  uint32_t n = (uint32_t)sc.lights.size();
  if (kinds) for (uint32_t i = 0; i < n; ++i) kinds[i] = sc.lights[i].kind;
  if (cdf) for (uint32_t i = 0; i <= n; ++i) cdf[i] = sc.light_cdf[i];
  return n;
}

}  // extern "C"
