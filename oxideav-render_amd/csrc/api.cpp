// Host runtime and C ABI of liboxr.so (include/oxr/pathtrace.h): the
// progressive PathTracer of oxideav-render (src/pathtrace.rs:2336-2735)
// with the scene, the BVH, the light tables and the accumulator resident
// on one HIP device, refined by the wavefront kernels of kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "bvh.h"
#include "dev_scene.h"
#include "oxr/pathtrace.h"
#include "runtime.h"

using namespace oxr;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPC(expr)                                                                                 \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return fail(OXR_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

constexpr float INF_F = std::numeric_limits<float>::infinity();

// ---------------------------------------------------------------------
// Constant tables, computed on the host once per process.
// ---------------------------------------------------------------------

// Sobol' generator matrices, dims 0-3: van der Corput and the Joe-Kuo
// primitive polynomials (s, a, m) = (1,0,[1]), (2,1,[1,3]), (3,1,[1,3,1])
// (src/pathtrace.rs:319-346).
void make_sobol(uint32_t out[4][32]) {
  for (int k = 0; k < 32; ++k) out[0][k] = 1u << (31 - k);
  struct P {
    int s;
    uint32_t a;
    uint32_t m[3];
  };
  const P polys[3] = {{1, 0, {1, 0, 0}}, {2, 1, {1, 3, 0}}, {3, 1, {1, 3, 1}}};
  for (int d = 0; d < 3; ++d) {
    uint32_t m[32];
    const P& p = polys[d];
    for (int k = 0; k < p.s; ++k) m[k] = p.m[k];
    for (int k = p.s; k < 32; ++k) {
      uint32_t v = m[k - p.s] ^ (m[k - p.s] << p.s);
      for (int j = 1; j < p.s; ++j)
        if ((p.a >> (p.s - 1 - j)) & 1u) v ^= m[k - j] << j;
      m[k] = v;
    }
    for (int k = 0; k < 32; ++k) out[d + 1][k] = m[k] << (31 - k);
  }
}

// Sheen directional-albedo table E(cos, roughness) by the reference's
// f32 midpoint quadrature (src/pathtrace.rs:645-674).
void make_sheen(float out[512]) {
  const float PI = 3.14159265358979323846f, HPI = 1.57079632679489661923f;
  auto dch = [&](float nh, float a) {
    float inv = 1.0f / std::fmax(a, 1.0e-3f);
    float s2 = std::fmax(1.0f - nh * nh, 0.0f);
    return (2.0f + inv) * std::pow(s2, inv * 0.5f) / (2.0f * PI);
  };
  auto vn = [](float nl, float nv) { return 1.0f / std::fmax(4.0f * (nl + nv - nl * nv), 1.0e-6f); };
  for (int ri = 0; ri < 16; ++ri) {
    float r = (float)ri / 15.0f;
    float a = std::fmax(r * r, 1.0e-3f);
    for (int ci = 0; ci < 32; ++ci) {
      float cv = std::fmax(((float)ci + 0.5f) / 32.0f, 1.0e-3f);
      float vx = std::sqrt(1.0f - cv * cv), vz = cv;
      float sum = 0.0f;
      for (int i = 0; i < 48; ++i) {
        float th = ((float)i + 0.5f) / 48.0f * HPI;
        float st = std::sin(th), ct = std::cos(th);
        for (int j = 0; j < 24; ++j) {
          float ph = ((float)j + 0.5f) / 24.0f * PI;
          float lx = st * std::cos(ph), ly = st * std::sin(ph), lz = ct;
          float hx = vx + lx, hy = 0.0f + ly, hz = vz + lz;
          float len = std::sqrt(hx * hx + hy * hy + hz * hz);
          float nhz = len < 1.1920929e-7f ? 0.0f : hz / len;
          float f = dch(nhz, a) * vn(ct, cv);
          sum += f * ct * st;
        }
      }
      float dw = (HPI / 48.0f) * (PI / 24.0f) * 2.0f;
      float v = sum * dw;
      if (v < 0.0f) v = 0.0f;
      if (v > 1.0f) v = 1.0f;
      out[ri * 32 + ci] = v;
    }
  }
}

float srgb_to_linear(float c) {
  if (c <= 0.04045f) return c / 12.92f;
  return std::pow((c + 0.055f) / 1.055f, 2.4f);
}

struct Tables {
  uint32_t sobol[4][32];
  float sheen[512];
  float srgb_th[255];
  Tables() {
    make_sobol(sobol);
    make_sheen(sheen);
    for (int k = 0; k < 255; ++k) srgb_th[k] = srgb_to_linear(((float)k + 0.5f) / 255.0f);
  }
};

const Tables& tables() {
  static const Tables t;
  return t;
}

std::mutex g_const_mu;
bool g_const_done[256];

int ensure_device(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(OXR_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= n || device >= 256) return fail(OXR_ERR_INVALID, "device %d out of range (%d)", device, n);
  HIPC(hipSetDevice(device));
  std::lock_guard<std::mutex> g(g_const_mu);
  if (!g_const_done[device]) {
    const Tables& t = tables();
    HIPC(launch::upload_constants(&t.sobol[0][0], t.sheen, t.srgb_th));
    g_const_done[device] = true;
  }
  return OXR_OK;
}

template <class T>
int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(OXR_ERR_MEMORY, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  }
  return OXR_OK;
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// PreparedMaterial::from_material (src/prepare.rs:623-668) + the
// untextured reads of TraceScene::material_oriented (src/trace.rs:724-922).
DevMaterial prepare_material(const oxr_material& m) {
  auto fin = [](float v, float d) { return std::isfinite(v) ? v : d; };
  auto cl = [](float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
  };
  DevMaterial d;
  std::memset(&d, 0, sizeof(d));
  for (int k = 0; k < 3; ++k) d.base[k] = fin(m.base_color[k], 1.0f);
  d.base[3] = cl(fin(m.base_color[3], 1.0f), 0.0f, 1.0f);
  d.metallic = cl(fin(m.metallic, 1.0f), 0.0f, 1.0f);
  d.roughness = cl(fin(m.roughness, 1.0f), 0.0f, 1.0f);
  d.ior = std::fmax(fin(m.ior, 1.5f), 1.0f);
  d.specular = cl(fin(m.specular, 1.0f), 0.0f, 1.0f);
  for (int k = 0; k < 3; ++k) {
    d.emissive[k] = std::fmax(fin(m.emissive[k], 0.0f), 0.0f);
    d.spec_color[k] = std::fmax(fin(m.specular_color[k], 1.0f), 0.0f);
    d.att_color[k] = cl(fin(m.attenuation_color[k], 1.0f), 0.0f, 1.0f);
    d.sheen[k] = cl(fin(m.sheen_color[k], 0.0f), 0.0f, 1.0f);
    d.dt_color[k] = cl(fin(m.diffuse_transmission_color[k], 1.0f), 0.0f, 1.0f);
  }
  d.flags = ((m.flags & OXR_MAT_DOUBLE_SIDED) ? MAT_DS : 0u) | ((m.flags & OXR_MAT_UNLIT) ? MAT_UNLIT : 0u);
  d.transmission = cl(fin(m.transmission, 0.0f), 0.0f, 1.0f);
  d.thickness = std::fmax(fin(m.thickness, 0.0f), 0.0f);
  float ad = m.attenuation_distance;
  d.att_dist = (std::isfinite(ad) && ad > 0.0f) ? ad : INF_F;
  d.clearcoat = cl(fin(m.clearcoat, 0.0f), 0.0f, 1.0f);
  d.cc_rough = cl(fin(m.clearcoat_roughness, 0.0f), 0.0f, 1.0f);
  d.sheen_rough = cl(fin(m.sheen_roughness, 0.0f), 0.0f, 1.0f);
  d.iridescence = cl(fin(m.iridescence, 0.0f), 0.0f, 1.0f);
  d.irid_ior = std::fmax(fin(m.iridescence_ior, 1.3f), 1.0f);
  d.irid_thick = std::fmax(fin(m.iridescence_thickness, 400.0f), 0.0f);
  d.dt = cl(fin(m.diffuse_transmission, 0.0f), 0.0f, 1.0f);
  d.dispersion = std::fmax(fin(m.dispersion, 0.0f), 0.0f);
  return d;
}

// The fallback material of primitives without one
// (PreparedMaterial::fallback, src/prepare.rs:600-621).
oxr_material fallback_material() {
  oxr_material m;
  std::memset(&m, 0, sizeof(m));
  m.base_color[0] = 0.7f;
  m.base_color[1] = 0.7f;
  m.base_color[2] = 0.75f;
  m.base_color[3] = 1.0f;
  m.metallic = 0.0f;
  m.roughness = 1.0f;
  m.ior = 1.5f;
  m.specular = 1.0f;
  for (int k = 0; k < 3; ++k) {
    m.specular_color[k] = 1.0f;
    m.attenuation_color[k] = 1.0f;
    m.diffuse_transmission_color[k] = 1.0f;
  }
  m.attenuation_distance = INF_F;
  m.iridescence_ior = 1.3f;
  m.iridescence_thickness = 400.0f;
  m.alpha_cutoff = 0.5f;
  return m;
}

// PreparedLight sanitising (src/prepare.rs:866-919) and the constants of
// PreparedLight::sample (src/prepare.rs:921-963). Returns false for a
// light that can never contribute (zero colour x intensity).
bool prepare_light(const oxr_light& l, DevLight& d) {
  auto fin0 = [](float v) { return std::isfinite(v) ? std::fmax(v, 0.0f) : 0.0f; };
  std::memset(&d, 0, sizeof(d));
  d.kind = l.kind <= OXR_LIGHT_SPOT ? l.kind : OXR_LIGHT_POINT;
  float inten = fin0(l.intensity);
  for (int k = 0; k < 3; ++k) {
    d.base[k] = fin0(l.color[k]) * inten;
    d.pos[k] = l.position[k];
    d.dir[k] = l.direction[k];
  }
  d.has_range = (d.kind != OXR_LIGHT_DIRECTIONAL && std::isfinite(l.range) && l.range > 0.0f) ? 1u : 0u;
  d.range = l.range;
  if (d.kind == OXR_LIGHT_SPOT) {
    const float HPI = 1.57079632679489661923f, QPI = 0.78539816339744830962f;
    float outer = QPI, inner = 0.0f;
    if (std::isfinite(l.outer_cone_angle)) outer = std::fmin(std::fmax(l.outer_cone_angle, 1.0e-4f), HPI);
    if (std::isfinite(l.inner_cone_angle)) inner = std::fmin(std::fmax(l.inner_cone_angle, 0.0f), outer);
    float cos_outer = std::cos(outer), cos_inner = std::cos(inner);
    float sc = 1.0f / std::fmax(cos_inner - cos_outer, 0.001f);
    d.spot_scale = sc;
    d.spot_offset = -cos_outer * sc;
  }
  return d.base[0] > 0.0f || d.base[1] > 0.0f || d.base[2] > 0.0f;
}

}  // namespace

// =====================================================================
// The tracer.
// =====================================================================
struct oxr_pt {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  // Scene.
  bool has_scene = false;
  DevScene sc{};
  Node4* d_nodes = nullptr;
  TriRec* d_tris = nullptr;
  float* d_pos = nullptr;
  float* d_nrm = nullptr;
  uint32_t* d_info = nullptr;
  DevMaterial* d_mats = nullptr;
  DevLight* d_lights = nullptr;
  uint32_t* d_ltris = nullptr;
  float* d_lpmf = nullptr;
  float* d_lcdf = nullptr;
  uint32_t* d_llookup = nullptr;
  oxr_bvh_info bvh{};
  uint32_t spill_cap = 4;
  // Frame.
  bool has_cam = false, has_opts = false;
  oxr_camera cam{};
  oxr_pt_options opts{};
  // Accumulator.
  double* d_accum = nullptr;
  double* d_sumsq = nullptr;
  size_t npx_alloc = 0;
  uint32_t samples = 0;
  // Wavefront buffers.
  PathBuffers pb{};
  Ctrl* d_ctrl = nullptr;
  uint32_t* d_spill = nullptr;
  size_t spill_alloc = 0;
  int grid_closest = 0, grid_any = 0;
  uint32_t path_budget = 1u << 21;
  // Scratch for the parity entry points.
  void* d_tmp = nullptr;
  size_t tmp_alloc = 0;
  // Stats / timing.
  bool timing = false;
  oxr_pt_stats stats{};
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Span {
    int cat;
    hipEvent_t a, b;
  };
  std::vector<Span> spans;

  ~oxr_pt() {
    if (device >= 0) (void)hipSetDevice(device);
    free_scene();
    free_paths();
    dfree(d_accum);
    dfree(d_sumsq);
    dfree(d_ctrl);
    if (d_tmp) (void)hipFree(d_tmp);
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void free_scene() {
    dfree(d_nodes);
    dfree(d_tris);
    dfree(d_pos);
    dfree(d_nrm);
    dfree(d_info);
    dfree(d_mats);
    dfree(d_lights);
    dfree(d_ltris);
    dfree(d_lpmf);
    dfree(d_lcdf);
    dfree(d_llookup);
    has_scene = false;
  }

  void free_paths() {
    for (int i = 0; i < 2; ++i) {
      dfree(pb.qray[i]);
      dfree(pb.qpid[i]);
    }
    dfree(pb.hits);
    dfree(pb.sray);
    dfree(pb.sidx);
    dfree(pb.beta);
    dfree(pb.rad);
    dfree(pb.prev);
    dfree(pb.sigma);
    dfree(pb.meta);
    dfree(pb.nee);
    dfree(pb.vis);
    dfree(pb.out);
    pb.cap = 0;
    pb.nslots = 0;
  }

  int tmp(size_t bytes, void** out) {
    if (bytes > tmp_alloc) {
      if (d_tmp) (void)hipFree(d_tmp);
      d_tmp = nullptr;
      tmp_alloc = 0;
      hipError_t e = hipMalloc(&d_tmp, bytes);
      if (e != hipSuccess) return fail(OXR_ERR_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
      tmp_alloc = bytes;
    }
    *out = d_tmp;
    return OXR_OK;
  }

  int ensure_paths(uint32_t cap, uint32_t nslots) {
    if (pb.cap >= cap && pb.nslots == nslots) return OXR_OK;
    free_paths();
    size_t c = cap, cs = (size_t)cap * nslots;
    int rc;
    for (int i = 0; i < 2; ++i) {
      if ((rc = dalloc(&pb.qray[i], 2 * c))) return rc;
      if ((rc = dalloc(&pb.qpid[i], c))) return rc;
    }
    if ((rc = dalloc(&pb.hits, c))) return rc;
    if ((rc = dalloc(&pb.sray, 2 * cs))) return rc;
    if ((rc = dalloc(&pb.sidx, cs))) return rc;
    if ((rc = dalloc(&pb.beta, c))) return rc;
    if ((rc = dalloc(&pb.rad, c))) return rc;
    if ((rc = dalloc(&pb.prev, c))) return rc;
    if ((rc = dalloc(&pb.sigma, c))) return rc;
    if ((rc = dalloc(&pb.meta, c))) return rc;
    if ((rc = dalloc(&pb.nee, cs))) return rc;
    if ((rc = dalloc(&pb.vis, cs))) return rc;
    if ((rc = dalloc(&pb.out, c))) return rc;
    HIPC(hipMemsetAsync(pb.vis, 0, cs, stream));
    pb.cap = cap;
    pb.nslots = nslots;
    return OXR_OK;
  }

  int ensure_spill() {
    int g = std::max(grid_closest, grid_any);
    size_t need = (size_t)g * launch::trace_block_threads() * spill_cap;
    if (need <= spill_alloc && d_spill) return OXR_OK;
    dfree(d_spill);
    spill_alloc = 0;
    int rc = dalloc(&d_spill, need);
    if (rc) return rc;
    spill_alloc = need;
    return OXR_OK;
  }

  int ensure_accum() {
    size_t npx = (size_t)std::max(opts.width, 1u) * std::max(opts.height, 1u);
    if (npx != npx_alloc) {
      dfree(d_accum);
      dfree(d_sumsq);
      npx_alloc = 0;
      int rc;
      if ((rc = dalloc(&d_accum, 4 * npx))) return rc;
      if ((rc = dalloc(&d_sumsq, 3 * npx))) return rc;
      npx_alloc = npx;
    }
    return OXR_OK;
  }

  int reset() {
    samples = 0;
    if (npx_alloc) {
      launch::fill_f64(stream, d_accum, 4 * npx_alloc);
      launch::fill_f64(stream, d_sumsq, 3 * npx_alloc);
      HIPC(hipGetLastError());
    }
    return OXR_OK;
  }

  void span_begin(int cat, hipEvent_t& a) {
    if (!timing) return;
    a = next_event();
    (void)hipEventRecord(a, stream);
    (void)cat;
  }
  void span_end(int cat, hipEvent_t a) {
    if (!timing) return;
    hipEvent_t b = next_event();
    (void)hipEventRecord(b, stream);
    spans.push_back({cat, a, b});
  }
  hipEvent_t next_event() {
    if (ev_used == ev_pool.size()) {
      hipEvent_t e;
      (void)hipEventCreate(&e);
      ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
  }

  DevFrame frame() const {
    DevFrame f;
    for (int k = 0; k < 3; ++k) {
      f.eye[k] = cam.eye[k];
      f.forward[k] = cam.forward[k];
      f.side[k] = cam.side[k];
      f.up[k] = cam.up[k];
    }
    f.half_w = cam.half_w;
    f.half_h = cam.half_h;
    f.projection = cam.projection;
    f.width = std::max(opts.width, 1u);
    f.height = std::max(opts.height, 1u);
    f.max_bounces = std::min(opts.max_bounces, 65534u);
    f.rr_start = opts.rr_start;
    f.seed = opts.seed;
    f.strategy = opts.strategy <= 2u ? opts.strategy : 0u;
    f.clamp = std::isfinite(opts.clamp) ? opts.clamp : 0.0f;
    f.ambient = std::isfinite(opts.ambient) ? std::fmax(opts.ambient, 0.0f) : 0.0f;
    f.nslots = sc.num_lights + 1;
    return f;
  }

  int refine(uint32_t n);
};

int oxr_pt::refine(uint32_t n) {
  if (!has_scene || !has_cam || !has_opts) return OXR_OK;  // no-op before sync
  if (n == 0) return OXR_OK;
  n = std::min(n, std::numeric_limits<uint32_t>::max() - samples);
  if (n == 0) return OXR_OK;
  int rc;
  HIPC(hipSetDevice(device));
  if ((rc = ensure_accum())) return rc;
  const uint32_t W = std::max(opts.width, 1u), H = std::max(opts.height, 1u);
  const uint64_t npx = (uint64_t)W * H;
  if (npx > (1ull << 31)) return fail(OXR_ERR_INVALID, "image too large (%llu pixels)", (unsigned long long)npx);
  DevFrame fr = frame();
  const uint32_t nslots = fr.nslots;
  // Batch = pixel range x sample range, at most path_budget paths.
  uint32_t npx_b = (uint32_t)std::min<uint64_t>(npx, path_budget);
  uint32_t S_b = npx_b < npx ? 1u : std::max(1u, std::min(n, (uint32_t)(path_budget / npx)));
  uint32_t cap = npx_b * S_b;
  if ((rc = ensure_paths(cap, nslots))) return rc;
  if ((rc = ensure_spill())) return rc;
  launch::zero_ctrl(stream, d_ctrl);
  spans.clear();
  ev_used = 0;
  const int grid_shade = (int)std::min<uint64_t>((cap + 255) / 256, (uint64_t)num_cu * 8);
  uint64_t paths = 0;
  for (uint32_t s0 = 0; s0 < n; s0 += S_b) {
    uint32_t S = std::min(S_b, n - s0);
    for (uint64_t px0 = 0; px0 < npx; px0 += npx_b) {
      uint32_t npx_c = (uint32_t)std::min<uint64_t>(npx_b, npx - px0);
      uint32_t P = npx_c * S;
      paths += P;
      hipEvent_t ea{};
      span_begin(3, ea);
      launch::camera(stream, fr, P, npx_c, (uint32_t)px0, samples + s0, pb, d_ctrl);
      span_end(3, ea);
      uint32_t cur = 0;
      for (uint32_t k = 0; k <= fr.max_bounces; ++k) {
        launch::prep(stream, d_ctrl, cur, k > 0 ? 1u : 0u);
        hipEvent_t e1{};
        span_begin(0, e1);
        launch::closest(stream, grid_closest, sc, true, pb.qray[cur], pb.qpid[cur], pb.meta, 0, &d_ctrl->qcount[cur],
                        &d_ctrl->work[0], 0, pb.hits, d_spill, spill_cap, &d_ctrl->error);
        span_end(0, e1);
        hipEvent_t e2{};
        span_begin(2, e2);
        launch::shade(stream, grid_shade, sc, fr, pb, d_ctrl, cur);
        span_end(2, e2);
        hipEvent_t e3{};
        span_begin(1, e3);
        launch::anyhit(stream, grid_any, sc, true, pb.sray, pb.sidx, pb.meta, nslots, 0, &d_ctrl->scount,
                       &d_ctrl->work[1], pb.vis, d_spill, spill_cap, &d_ctrl->error);
        span_end(1, e3);
        cur ^= 1u;
        // Long paths: stop launching once every path has ended.
        if (k >= 8 && (k % 4) == 0) {
          uint32_t left = 0;
          HIPC(hipMemcpyAsync(&left, &d_ctrl->qcount[cur], sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
          HIPC(hipStreamSynchronize(stream));
          if (left == 0) break;
        }
      }
      launch::prep(stream, d_ctrl, cur, 1u);
      hipEvent_t e4{};
      span_begin(3, e4);
      launch::finalize(stream, P, nslots, pb);
      launch::reduce(stream, npx_c, S, pb.out, d_accum + 4 * px0, d_sumsq + 3 * px0);
      span_end(3, e4);
    }
  }
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(stream));
  Ctrl c{};
  HIPC(hipMemcpy(&c, d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
  if (c.error) return fail(OXR_ERR_DEVICE, "traversal stack overflow (BVH depth %u)", bvh.max_depth);
  stats.paths = paths;
  stats.closest_rays = c.rays_closest;
  stats.shadow_rays = c.rays_shadow;
  stats.ms_closest = stats.ms_anyhit = stats.ms_shade = stats.ms_other = 0.0;
  stats.launches_closest = stats.launches_anyhit = 0;
  for (auto& s : spans) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, s.a, s.b);
    if (s.cat == 0) {
      stats.ms_closest += ms;
      stats.launches_closest++;
    } else if (s.cat == 1) {
      stats.ms_anyhit += ms;
      stats.launches_anyhit++;
    } else if (s.cat == 2) {
      stats.ms_shade += ms;
    } else {
      stats.ms_other += ms;
    }
  }
  samples += n;
  return OXR_OK;
}

// =====================================================================
// C ABI.
// =====================================================================
extern "C" {

const char* oxr_last_error(void) { return g_err.c_str(); }

int oxr_device_count(int* out) {
  if (!out) return fail(OXR_ERR_INVALID, "null out");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *out = n;
  return OXR_OK;
}

int oxr_pt_create(int device, oxr_pt** out) {
  if (!out) return fail(OXR_ERR_INVALID, "null out");
  *out = nullptr;
  int rc = ensure_device(device);
  if (rc) return rc;
  oxr_pt* h = new oxr_pt();
  h->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess) {
    delete h;
    return fail(OXR_ERR_DEVICE, "device %d: stream creation failed", device);
  }
  h->num_cu = prop.multiProcessorCount;
  h->grid_closest = h->num_cu * launch::trace_blocks_per_cu(false);
  h->grid_any = h->num_cu * launch::trace_blocks_per_cu(true);
  if (const char* b = std::getenv("OXR_PATH_BUDGET")) {
    long v = std::atol(b);
    if (v >= 4096 && v <= (1l << 26)) h->path_budget = (uint32_t)v;
  }
  if ((rc = dalloc(&h->d_ctrl, 1))) {
    delete h;
    return rc;
  }
  *out = h;
  return OXR_OK;
}

int oxr_pt_destroy(oxr_pt* h) {
  delete h;
  return OXR_OK;
}

int oxr_pt_upload_scene(oxr_pt* h, const oxr_scene_desc* d) {
  if (!h || !d) return fail(OXR_ERR_INVALID, "null argument");
  if (d->num_tris > 0 && (!d->positions || !d->material_ids))
    return fail(OXR_ERR_INVALID, "positions and material_ids are required");
  if (d->num_tris >= (1u << 31)) return fail(OXR_ERR_INVALID, "too many triangles (%u)", d->num_tris);
  if (d->num_materials > 0 && !d->materials) return fail(OXR_ERR_INVALID, "null materials");
  if (d->num_lights > 0 && !d->lights) return fail(OXR_ERR_INVALID, "null lights");
  HIPC(hipSetDevice(h->device));
  HIPC(hipStreamSynchronize(h->stream));
  h->free_scene();
  const uint32_t n = d->num_tris;
  // Materials (+ the fallback for out-of-range ids).
  std::vector<DevMaterial> mats;
  for (uint32_t i = 0; i < d->num_materials; ++i) mats.push_back(prepare_material(d->materials[i]));
  const uint32_t fallback = (uint32_t)mats.size();
  mats.push_back(prepare_material(fallback_material()));
  std::vector<uint32_t> info(std::max(n, 1u), 0), tflags(std::max(n, 1u), 0);
  std::vector<float> alpha(std::max(n, 1u), 1.0f);
  std::vector<oxr_material> src(d->materials, d->materials + d->num_materials);
  src.push_back(fallback_material());
  for (uint32_t g = 0; g < n; ++g) {
    uint32_t mid = d->material_ids[g];
    if (mid >= d->num_materials) mid = fallback;
    bool has_n = d->normals && (d->tri_flags ? (d->tri_flags[g] & OXR_TRI_HAS_NORMALS) != 0 : true);
    info[g] = mid | (has_n ? TI_HAS_NORMALS : 0u);
    const DevMaterial& m = mats[mid];
    uint32_t f = (m.flags & MAT_DS) ? TRI_DOUBLE_SIDED : 0u;
    if (src[mid].alpha_mode == OXR_ALPHA_MASK && !(m.base[3] >= src[mid].alpha_cutoff)) f |= TRI_MASK_REJECT;
    if (src[mid].alpha_mode == OXR_ALPHA_BLEND) f |= TRI_BLEND;
    tflags[g] = f;
    alpha[g] = m.base[3];
  }
  // Punctual lights (zero-power lights contribute exactly nothing and are
  // dropped).
  std::vector<DevLight> lights;
  for (uint32_t i = 0; i < d->num_lights; ++i) {
    DevLight dl;
    if (prepare_light(d->lights[i], dl)) lights.push_back(dl);
  }
  if (lights.size() > 4096) return fail(OXR_ERR_INVALID, "too many punctual lights (%zu)", lights.size());
  // EmissiveLights::build (src/pathtrace.rs:1479-1543).
  std::vector<uint32_t> ltris, lookup(std::max(n, 1u), 0xffffffffu);
  std::vector<float> lpow;
  for (uint32_t g = 0; g < n; ++g) {
    const DevMaterial& m = mats[info[g] & TI_MAT_MASK];
    float e0 = (m.flags & MAT_UNLIT) ? 0.0f : m.emissive[0];
    float e1 = (m.flags & MAT_UNLIT) ? 0.0f : m.emissive[1];
    float e2 = (m.flags & MAT_UNLIT) ? 0.0f : m.emissive[2];
    float lum = 0.2126f * e0 + 0.7152f * e1 + 0.0722f * e2;
    if (std::isnan(lum) || lum <= 0.0f) continue;
    const float* p = d->positions + 9ull * g;
    float ax = p[3] - p[0], ay = p[4] - p[1], az = p[5] - p[2];
    float bx = p[6] - p[0], by = p[7] - p[1], bz = p[8] - p[2];
    float cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    float area = 0.5f * std::sqrt(cx * cx + cy * cy + cz * cz);
    if (std::isnan(area) || area <= 0.0f || !std::isfinite(area)) continue;
    float sides = (m.flags & MAT_DS) ? 2.0f : 1.0f;
    lookup[g] = (uint32_t)ltris.size();
    ltris.push_back(g);
    lpow.push_back(area * lum * sides);
  }
  double total = 0.0;
  for (float p : lpow) total += (double)p;
  std::vector<float> pmf, cdf;
  double acc = 0.0;
  for (float pw : lpow) {
    pmf.push_back((float)((double)pw / total));
    acc += (double)pw / total;
    cdf.push_back((float)acc);
  }
  if (!cdf.empty()) cdf.back() = 1.0f;
  // BVH over the soup.
  Bvh4 bvh = build_bvh4(d->positions, n, tflags.data(), alpha.data(), 4, 0);
  if (bvh.nodes.empty()) {
    Node4 root;
    for (int i = 0; i < 4; ++i) {
      root.lo_x[i] = root.lo_y[i] = root.lo_z[i] = INF_F;
      root.hi_x[i] = root.hi_y[i] = root.hi_z[i] = -INF_F;
      root.child[i] = NODE_EMPTY;
      root.count[i] = 0;
    }
    bvh.nodes.push_back(root);
  }
  int rc;
  auto up = [&](auto** dst, const auto* srcp, size_t count) -> int {
    int r = dalloc(dst, count);
    if (r) return r;
    if (count) {
      hipError_t e = hipMemcpy(*dst, srcp, count * sizeof(**dst), hipMemcpyHostToDevice);
      if (e != hipSuccess) return fail(OXR_ERR_DEVICE, "upload: %s", hipGetErrorString(e));
    }
    return OXR_OK;
  };
  if ((rc = up(&h->d_nodes, bvh.nodes.data(), bvh.nodes.size()))) return rc;
  if ((rc = up(&h->d_tris, bvh.tris.data(), bvh.tris.size()))) return rc;
  if ((rc = up(&h->d_pos, d->positions, 9ull * n))) return rc;
  if (d->normals) {
    if ((rc = up(&h->d_nrm, d->normals, 9ull * n))) return rc;
  }
  if ((rc = up(&h->d_info, info.data(), info.size()))) return rc;
  if ((rc = up(&h->d_mats, mats.data(), mats.size()))) return rc;
  if ((rc = up(&h->d_lights, lights.data(), lights.size()))) return rc;
  if ((rc = up(&h->d_ltris, ltris.data(), ltris.size()))) return rc;
  if ((rc = up(&h->d_lpmf, pmf.data(), pmf.size()))) return rc;
  if ((rc = up(&h->d_lcdf, cdf.data(), cdf.size()))) return rc;
  if ((rc = up(&h->d_llookup, lookup.data(), lookup.size()))) return rc;
  DevScene& sc = h->sc;
  sc.nodes = h->d_nodes;
  sc.tris = h->d_tris;
  sc.pos = h->d_pos;
  sc.nrm = h->d_nrm;
  sc.tri_info = h->d_info;
  sc.mats = h->d_mats;
  sc.lights = h->d_lights;
  sc.l_tris = h->d_ltris;
  sc.l_pmf = h->d_lpmf;
  sc.l_cdf = h->d_lcdf;
  sc.l_lookup = h->d_llookup;
  sc.num_tris = n;
  sc.num_mats = (uint32_t)mats.size();
  sc.num_lights = (uint32_t)lights.size();
  sc.num_emissive = (uint32_t)ltris.size();
  sc.has_blend = 0;
  for (uint32_t g = 0; g < n; ++g)
    if (tflags[g] & TRI_BLEND) sc.has_blend = 1;
  h->bvh.num_nodes = (uint32_t)bvh.nodes.size();
  h->bvh.num_leaves = bvh.num_leaves;
  h->bvh.num_tris = (uint32_t)bvh.tris.size();
  h->bvh.max_depth = bvh.max_depth;
  h->bvh.sah_cost = bvh.sah_cost;
  h->bvh.build_ms = bvh.build_ms;
  // Stack bound: each level leaves at most 3 siblings behind.
  int lds = launch::stack_lds_entries();
  h->spill_cap = (uint32_t)std::max(4, 4 * (int)bvh.max_depth + 8 - lds);
  h->spill_alloc = 0;
  dfree(h->d_spill);
  h->has_scene = true;
  return h->reset();
}

int oxr_pt_set_camera(oxr_pt* h, const oxr_camera* cam, int* reset_out) {
  if (!h || !cam) return fail(OXR_ERR_INVALID, "null argument");
  bool changed = !h->has_cam || std::memcmp(&h->cam, cam, sizeof(oxr_camera)) != 0;
  h->cam = *cam;
  h->has_cam = true;
  if (reset_out) *reset_out = changed ? 1 : 0;
  if (changed) return h->reset();
  return OXR_OK;
}

int oxr_pt_set_options(oxr_pt* h, const oxr_pt_options* o, int* reset_out) {
  if (!h || !o) return fail(OXR_ERR_INVALID, "null argument");
  // radiance_key (src/pathtrace.rs:2402-2418): everything except the
  // sample target can change the estimate.
  oxr_pt_options a = h->opts, b = *o;
  a.samples_per_pixel = b.samples_per_pixel = 0;
  bool changed = !h->has_opts || std::memcmp(&a, &b, sizeof(a)) != 0;
  h->opts = *o;
  h->has_opts = true;
  if (reset_out) *reset_out = changed ? 1 : 0;
  if (changed) {
    HIPC(hipSetDevice(h->device));
    int rc = h->ensure_accum();
    if (rc) return rc;
    return h->reset();
  }
  return OXR_OK;
}

int oxr_pt_reset(oxr_pt* h) {
  if (!h) return fail(OXR_ERR_INVALID, "null handle");
  HIPC(hipSetDevice(h->device));
  return h->reset();
}

int oxr_pt_refine(oxr_pt* h, uint32_t n) {
  if (!h) return fail(OXR_ERR_INVALID, "null handle");
  return h->refine(n);
}

int oxr_pt_samples(oxr_pt* h, uint32_t* out) {
  if (!h || !out) return fail(OXR_ERR_INVALID, "null argument");
  *out = h->samples;
  return OXR_OK;
}

int oxr_pt_read_accum(oxr_pt* h, double* accum, double* sum_sq) {
  if (!h || !accum) return fail(OXR_ERR_INVALID, "null argument");
  size_t npx = (size_t)std::max(h->opts.width, 1u) * std::max(h->opts.height, 1u);
  if (!h->npx_alloc) {
    std::memset(accum, 0, npx * 4 * sizeof(double));
    if (sum_sq) std::memset(sum_sq, 0, npx * 3 * sizeof(double));
    return OXR_OK;
  }
  HIPC(hipSetDevice(h->device));
  HIPC(hipStreamSynchronize(h->stream));
  HIPC(hipMemcpy(accum, h->d_accum, h->npx_alloc * 4 * sizeof(double), hipMemcpyDeviceToHost));
  if (sum_sq) HIPC(hipMemcpy(sum_sq, h->d_sumsq, h->npx_alloc * 3 * sizeof(double), hipMemcpyDeviceToHost));
  return OXR_OK;
}

static int resolve_common(oxr_pt* h, const oxr_display* disp, bool rgba8, void* out) {
  if (!h || !disp || !out) return fail(OXR_ERR_INVALID, "null argument");
  size_t npx = (size_t)std::max(h->opts.width, 1u) * std::max(h->opts.height, 1u);
  float bg[4];
  for (int k = 0; k < 3; ++k) bg[k] = srgb_to_linear((float)disp->background[k] / 255.0f);
  bg[3] = (float)disp->background[3] / 255.0f;
  HIPC(hipSetDevice(h->device));
  int rc;
  if ((rc = h->ensure_accum())) return rc;
  void* d_out;
  size_t bytes = npx * (rgba8 ? 4 : 16);
  if ((rc = h->tmp(bytes, &d_out))) return rc;
  launch::resolve(h->stream, rgba8, (uint32_t)npx, h->samples, h->d_accum, bg, disp->background, disp->tone_map,
                  disp->exposure, d_out);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(out, d_out, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  return OXR_OK;
}

int oxr_pt_resolve_hdr(oxr_pt* h, const oxr_display* disp, float* out) { return resolve_common(h, disp, false, out); }
int oxr_pt_resolve_rgba8(oxr_pt* h, const oxr_display* disp, uint8_t* out) {
  return resolve_common(h, disp, true, out);
}

int oxr_pt_set_timing(oxr_pt* h, int enable) {
  if (!h) return fail(OXR_ERR_INVALID, "null handle");
  h->timing = enable != 0;
  return OXR_OK;
}

int oxr_pt_stats_get(oxr_pt* h, oxr_pt_stats* out) {
  if (!h || !out) return fail(OXR_ERR_INVALID, "null argument");
  *out = h->stats;
  return OXR_OK;
}

int oxr_pt_bvh_info(oxr_pt* h, oxr_bvh_info* out) {
  if (!h || !out) return fail(OXR_ERR_INVALID, "null argument");
  if (!h->has_scene) return fail(OXR_ERR_STATE, "no scene uploaded");
  *out = h->bvh;
  return OXR_OK;
}

int oxr_pt_bvh_export(oxr_pt* h, void* nodes, void* tris) {
  if (!h || !nodes || !tris) return fail(OXR_ERR_INVALID, "null argument");
  if (!h->has_scene) return fail(OXR_ERR_STATE, "no scene uploaded");
  HIPC(hipSetDevice(h->device));
  HIPC(hipStreamSynchronize(h->stream));
  HIPC(hipMemcpy(nodes, h->d_nodes, (size_t)h->bvh.num_nodes * sizeof(Node4), hipMemcpyDeviceToHost));
  if (h->bvh.num_tris)
    HIPC(hipMemcpy(tris, h->d_tris, (size_t)h->bvh.num_tris * sizeof(TriRec), hipMemcpyDeviceToHost));
  return OXR_OK;
}

int oxr_sobol_owen_4d(int device, uint32_t n, const uint32_t* index, const uint32_t* pattern_seed, float* out) {
  if (n == 0) return OXR_OK;
  if (!index || !pattern_seed || !out) return fail(OXR_ERR_INVALID, "null argument");
  int rc = ensure_device(device);
  if (rc) return rc;
  uint32_t *d_i = nullptr, *d_s = nullptr;
  float4* d_o = nullptr;
  if ((rc = dalloc(&d_i, n)) || (rc = dalloc(&d_s, n)) || (rc = dalloc(&d_o, n))) {
    dfree(d_i);
    dfree(d_s);
    dfree(d_o);
    return rc;
  }
  hipError_t e = hipMemcpy(d_i, index, n * 4ull, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_s, pattern_seed, n * 4ull, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch::sobol(nullptr, n, d_i, d_s, d_o);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, d_o, n * 16ull, hipMemcpyDeviceToHost);
  dfree(d_i);
  dfree(d_s);
  dfree(d_o);
  if (e != hipSuccess) return fail(OXR_ERR_DEVICE, "sobol: %s", hipGetErrorString(e));
  return OXR_OK;
}

static int trace_host(oxr_pt* h, uint32_t n, const oxr_ray* rays, uint32_t flags, oxr_hit* hits, uint8_t* occ) {
  if (!h || (n && !rays)) return fail(OXR_ERR_INVALID, "null argument");
  if (!h->has_scene) return fail(OXR_ERR_STATE, "no scene uploaded");
  if (n == 0) return OXR_OK;
  HIPC(hipSetDevice(h->device));
  int rc;
  if ((rc = h->ensure_spill())) return rc;
  size_t rbytes = (size_t)n * 32, obytes = occ ? (size_t)n : (size_t)n * 16;
  size_t off = (rbytes + 255) & ~(size_t)255;
  void* buf;
  if ((rc = h->tmp(off + obytes + 256, &buf))) return rc;
  float4* d_rays = (float4*)buf;
  void* d_out = (char*)buf + off;
  HIPC(hipMemcpyAsync(d_rays, rays, rbytes, hipMemcpyHostToDevice, h->stream));
  launch::zero_ctrl(h->stream, h->d_ctrl);
  if (occ)
    launch::anyhit(h->stream, h->grid_any, h->sc, false, d_rays, nullptr, nullptr, 1, n, nullptr, &h->d_ctrl->work[1],
                   (uint8_t*)d_out, h->d_spill, h->spill_cap, &h->d_ctrl->error);
  else
    launch::closest(h->stream, h->grid_closest, h->sc, false, d_rays, nullptr, nullptr, n, nullptr, &h->d_ctrl->work[0],
                    flags & 1u, (uint4*)d_out, h->d_spill, h->spill_cap, &h->d_ctrl->error);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(occ ? (void*)occ : (void*)hits, d_out, obytes, hipMemcpyDeviceToHost, h->stream));
  Ctrl c{};
  HIPC(hipMemcpyAsync(&c, h->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, h->stream));
  HIPC(hipStreamSynchronize(h->stream));
  if (c.error) return fail(OXR_ERR_DEVICE, "traversal stack overflow");
  return OXR_OK;
}

int oxr_trace_closest(oxr_pt* h, uint32_t n, const oxr_ray* rays, uint32_t flags, oxr_hit* hits) {
  if (n && !hits) return fail(OXR_ERR_INVALID, "null hits");
  return trace_host(h, n, rays, flags, hits, nullptr);
}

int oxr_trace_any(oxr_pt* h, uint32_t n, const oxr_ray* rays, uint8_t* occluded) {
  if (n && !occluded) return fail(OXR_ERR_INVALID, "null output");
  return trace_host(h, n, rays, 0, nullptr, occluded);
}

int oxr_trace_closest_device(oxr_pt* h, uint32_t n, const void* d_rays, void* d_hits, void* stream) {
  if (!h || (n && (!d_rays || !d_hits))) return fail(OXR_ERR_INVALID, "null argument");
  if (!h->has_scene) return fail(OXR_ERR_STATE, "no scene uploaded");
  if (n == 0) return OXR_OK;
  HIPC(hipSetDevice(h->device));
  int rc;
  if ((rc = h->ensure_spill())) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  launch::zero_ctrl(s, h->d_ctrl);
  launch::closest(s, h->grid_closest, h->sc, false, (const float4*)d_rays, nullptr, nullptr, n, nullptr,
                  &h->d_ctrl->work[0], 0, (uint4*)d_hits, h->d_spill, h->spill_cap, &h->d_ctrl->error);
  HIPC(hipGetLastError());
  return OXR_OK;
}

static int bsdf_common(int device, uint32_t n, int do_sample, const oxr_material_sample* ms, const float* v,
                       const float* ng, const uint8_t* front, const float* in, float* l, float* f, float* pdf,
                       uint8_t* ok) {
  if (n == 0) return OXR_OK;
  if (!ms || !v || !ng || !front || !in || !f || !pdf || (do_sample && (!l || !ok)))
    return fail(OXR_ERR_INVALID, "null argument");
  static_assert(sizeof(oxr_material_sample) == 35 * 4, "material sample is 35 floats");
  int rc = ensure_device(device);
  if (rc) return rc;
  size_t bytes = (size_t)n * (140 + 12 + 12 + 1 + 12 + 12 + 12 + 4 + 1) + 16 * 9;
  char* buf = nullptr;
  if ((rc = dalloc(&buf, bytes))) return rc;
  size_t o = 0;
  auto carve = [&](size_t b) {
    char* p = buf + o;
    o += (b + 15) & ~(size_t)15;
    return p;
  };
  float* d_ms = (float*)carve(n * 140ull);
  float* d_v = (float*)carve(n * 12ull);
  float* d_ng = (float*)carve(n * 12ull);
  uint8_t* d_fr = (uint8_t*)carve(n);
  float* d_in = (float*)carve(n * 12ull);
  float* d_l = (float*)carve(n * 12ull);
  float* d_f = (float*)carve(n * 12ull);
  float* d_pdf = (float*)carve(n * 4ull);
  uint8_t* d_ok = (uint8_t*)carve(n);
  hipError_t e = hipMemcpy(d_ms, ms, n * 140ull, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_v, v, n * 12ull, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_ng, ng, n * 12ull, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_fr, front, n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_in, in, n * 12ull, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch::bsdf(nullptr, n, do_sample, d_ms, d_v, d_ng, d_fr, d_in, d_l, d_f, d_pdf, d_ok);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(f, d_f, n * 12ull, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(pdf, d_pdf, n * 4ull, hipMemcpyDeviceToHost);
  if (e == hipSuccess && do_sample) e = hipMemcpy(l, d_l, n * 12ull, hipMemcpyDeviceToHost);
  if (e == hipSuccess && do_sample) e = hipMemcpy(ok, d_ok, n, hipMemcpyDeviceToHost);
  dfree(buf);
  if (e != hipSuccess) return fail(OXR_ERR_DEVICE, "bsdf: %s", hipGetErrorString(e));
  return OXR_OK;
}

int oxr_bsdf_eval(int device, uint32_t n, const oxr_material_sample* ms, const float* v, const float* ng,
                  const uint8_t* front, const float* l, float* f, float* pdf) {
  return bsdf_common(device, n, 0, ms, v, ng, front, l, nullptr, f, pdf, nullptr);
}

int oxr_bsdf_sample(int device, uint32_t n, const oxr_material_sample* ms, const float* v, const float* ng,
                    const uint8_t* front, const float* u, float* l, float* f, float* pdf, uint8_t* ok) {
  return bsdf_common(device, n, 1, ms, v, ng, front, u, l, f, pdf, ok);
}

}  // extern "C"
