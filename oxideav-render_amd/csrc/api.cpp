// Host runtime and C ABI of liboxr.so (include/oxr/pathtrace.h): the
// progressive PathTracer of oxideav-render (src/pathtrace.rs:2336-2735)
// with the scene, the BVH, the light tables and the accumulator resident
// on one HIP device, refined by the wavefront kernels of kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <chrono>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "bvh.h"
#include "dev_guide.h"
#include "dev_scene.h"
#include "guide_rt.h"
#include "oxr/guide.h"
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

// Fixed scratch of oxr_trace_closest_device_noalloc: its control block
// (work counter, overflow flag) and its traversal-stack spill, so that
// call allocates nothing. Every GPU the library runs on holds one copy
// until the process ends; calls on one GPU share it (overlapping calls
// race). g_scratch[d] is the device address of device d's copy: host code
// uses it, never the array's name, which in host code is only a stand-in.
constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
void* g_scratch[256];

int ensure_device(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(OXR_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= n || device >= 256) return fail(OXR_ERR_INVALID, "device %d out of range (%d)", device, n);
  HIPC(hipSetDevice(device));
  std::lock_guard<std::mutex> g(g_const_mu);
  if (!g_const_done[device]) {
    const Tables& t = tables();
    HIPC(launch::upload_constants(&t.sobol[0][0], t.sheen, t.srgb_th));
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(OXR_ERR_DEVICE, "device %d: address of the fixed scratch: %s", device, hipGetErrorString(e));
    g_scratch[device] = p;
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
  d.aniso = cl(fin(m.anisotropy, 0.0f), 0.0f, 1.0f);
  d.aniso_rot = fin(m.anisotropy_rotation, 0.0f);
  return d;
}

// ---------------------------------------------------------------------
// Textures: the mip chains of src/texture.rs:323-381 (level 0 decoded
// through the sRGB byte table or divided by 255; float texels as given;
// then 2x2 box filters with edge clamping down to 1x1), built once per
// (texture, colour space) a material uses.
// ---------------------------------------------------------------------
struct TexBuilder {
  const oxr_scene_desc* d;
  std::vector<float4> pool;
  std::vector<DevTexLevel> levels;
  std::vector<DevTexChain> chains;
  std::vector<int32_t> chain_of;  // [texture * 2 + linear] -> chain, -1 = not built

  int validate() {
    if (d->num_textures > 0 && !d->textures) return fail(OXR_ERR_INVALID, "null textures");
    for (uint32_t i = 0; i < d->num_textures; ++i) {
      const oxr_texture& t = d->textures[i];
      if (t.width == 0 || t.height == 0 || !t.pixels)
        return fail(OXR_ERR_INVALID, "texture %u: empty image", i);
      if (t.format > OXR_TEXFMT_RGBA32F || t.mag_filter > OXR_FILTER_LINEAR || t.min_filter > OXR_MIN_LINEAR_MIP_LINEAR ||
          t.wrap_s > OXR_WRAP_MIRRORED_REPEAT || t.wrap_t > OXR_WRAP_MIRRORED_REPEAT)
        return fail(OXR_ERR_INVALID, "texture %u: unknown format, filter or wrap mode", i);
    }
    chain_of.assign(2ull * d->num_textures, -1);
    return OXR_OK;
  }

  uint32_t chain(uint32_t tex, bool srgb) {
    int32_t& c = chain_of[2ull * tex + (srgb ? 0 : 1)];
    if (c >= 0) return (uint32_t)c;
    static const std::array<float, 256> lut = [] {
      std::array<float, 256> t{};
      for (int i = 0; i < 256; ++i) t[i] = srgb_to_linear((float)i / 255.0f);
      return t;
    }();
    const oxr_texture& t = d->textures[tex];
    DevTexChain ch{};
    ch.level0 = (uint32_t)levels.size();
    ch.w0 = t.width;
    ch.h0 = t.height;
    ch.mag_linear = t.mag_filter == OXR_FILTER_LINEAR ? 1u : 0u;
    ch.min_filter = t.min_filter;
    ch.wrap_s = t.wrap_s;
    ch.wrap_t = t.wrap_t;
    size_t n0 = (size_t)t.width * t.height;
    size_t base = pool.size();
    pool.resize(base + n0);
    if (t.format == OXR_TEXFMT_RGBA8) {
      const uint8_t* p = static_cast<const uint8_t*>(t.pixels);
      for (size_t i = 0; i < n0; ++i) {
        const uint8_t* q = p + 4 * i;
        pool[base + i] = srgb ? make_float4(lut[q[0]], lut[q[1]], lut[q[2]], (float)q[3] / 255.0f)
                              : make_float4((float)q[0] / 255.0f, (float)q[1] / 255.0f, (float)q[2] / 255.0f,
                                            (float)q[3] / 255.0f);
      }
    } else {
      const float* p = static_cast<const float*>(t.pixels);
      for (size_t i = 0; i < n0; ++i) pool[base + i] = make_float4(p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]);
    }
    levels.push_back(DevTexLevel{(uint64_t)base, t.width, t.height});
    while (!(levels.back().w == 1 && levels.back().h == 1)) {
      const DevTexLevel pv = levels.back();
      uint32_t w = std::max(pv.w / 2, 1u), h = std::max(pv.h / 2, 1u);
      size_t nb = pool.size();
      pool.resize(nb + (size_t)w * h);
      for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
          size_t x0 = std::min(2 * x, pv.w - 1), x1 = std::min(2 * x + 1, pv.w - 1);
          size_t y0 = std::min(2 * y, pv.h - 1), y1 = std::min(2 * y + 1, pv.h - 1);
          const float4* src = pool.data() + pv.off;
          float a[4] = {0, 0, 0, 0};
          for (size_t idx : {y0 * pv.w + x0, y0 * pv.w + x1, y1 * pv.w + x0, y1 * pv.w + x1}) {
            a[0] += src[idx].x;
            a[1] += src[idx].y;
            a[2] += src[idx].z;
            a[3] += src[idx].w;
          }
          pool[nb + (size_t)y * w + x] = make_float4(a[0] * 0.25f, a[1] * 0.25f, a[2] * 0.25f, a[3] * 0.25f);
        }
      levels.push_back(DevTexLevel{(uint64_t)nb, w, h});
    }
    ch.nlevels = (uint32_t)levels.size() - ch.level0;
    c = (int32_t)chains.size();
    chains.push_back(ch);
    return (uint32_t)c;
  }
  // The 1x1 level of a chain.
  float4 top(uint32_t c) const { return pool[levels[chains[c].level0 + chains[c].nlevels - 1].off]; }
};

// Colour space of each oxr_material_textures slot (src/trace.rs:717-895).
static bool slot_srgb(uint32_t slot) {
  return slot == TS_BASE || slot == TS_EMISSIVE || slot == TS_SPEC_COLOR || slot == TS_SHEEN_COLOR ||
         slot == TS_DT_COLOR;
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
// Scenes resident on the device.
// =====================================================================
// A scene prepared on the host, ready to upload: everything
// SceneBuf::upload_host copies to the device. The triangle arrays either
// point into the caller's scene description (a direct upload) or into the
// object's own copies (own(): a prebuilt scene that outlives the call).
struct HostScene {
  uint32_t n = 0;
  float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
  std::vector<DevMaterial> mats;
  std::vector<float4> tex_pool;
  std::vector<DevTexLevel> tex_levels;
  std::vector<DevTexChain> tex_chains;
  std::vector<int32_t> chain_of;
  std::vector<DevTexBind> binds;
  std::vector<uint2> mattex;
  bool any_attr = false;
  std::vector<uint32_t> attr;
  const float* attr_src[4] = {nullptr, nullptr, nullptr, nullptr};  // uv0, uv1, tangents, colors
  const float* pos = nullptr;
  const float* nrm = nullptr;
  std::vector<float> own_arrays[6];  // pos, nrm, uv0, uv1, tangents, colors
  std::vector<uint32_t> info;
  bool has_blend = false, has_alpha_tex = false;
  std::vector<DevLight> lights;
  std::vector<uint32_t> ltris, lookup;
  std::vector<float> pmf, cdf;
  std::vector<Node4Q> qnodes;
  std::vector<TriRec> tris;
  uint32_t num_nodes = 0, num_leaves = 0, max_depth = 0;
  double sah_cost = 0.0, build_ms = 0.0;

  static constexpr size_t ATTR_W[4] = {6, 6, 12, 12};  // floats per triangle

  // Copies the caller's triangle arrays, so the object outlives the call.
  void own() {
    auto keep = [&](const float*& src, std::vector<float>& dst, size_t count) {
      if (!src) return;
      dst.assign(src, src + count);
      src = dst.data();
    };
    keep(pos, own_arrays[0], 9ull * n);
    keep(nrm, own_arrays[1], 9ull * n);
    for (int k = 0; k < 4; ++k) keep(attr_src[k], own_arrays[2 + k], ATTR_W[k] * n);
  }
  size_t device_bytes() const {
    size_t b = qnodes.size() * sizeof(Node4Q) + tris.size() * sizeof(TriRec) + 36ull * n * (nrm ? 2 : 1) +
               info.size() * 4 + mats.size() * sizeof(DevMaterial) + lights.size() * sizeof(DevLight) +
               (ltris.size() + pmf.size() + cdf.size() + lookup.size()) * 4 + tex_pool.size() * sizeof(float4) +
               attr.size() * 4;
    for (int k = 0; k < 4; ++k)
      if (attr_src[k]) b += ATTR_W[k] * 4ull * n;
    return b;
  }
};
constexpr size_t HostScene::ATTR_W[4];

int prepare_scene(const oxr_scene_desc* d, bool fast, HostScene& hs);

struct SceneBuf {
  DevScene sc{};
  Node4Q* d_nodes = nullptr;
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
  // Textures and per-triangle attributes.
  DevTexChain* d_chains = nullptr;
  DevTexLevel* d_levels = nullptr;
  float4* d_texels = nullptr;
  DevTexBind* d_binds = nullptr;
  uint2* d_mattex = nullptr;
  float* d_uv[2] = {nullptr, nullptr};
  float* d_tan = nullptr;
  float* d_col = nullptr;
  uint32_t* d_attr = nullptr;
  DevScene* d_self = nullptr;  // a device copy of `sc` (the textured-alpha traversal reads it)
  std::vector<int32_t> chain_of;  // [texture * 2 + linear] -> chain (parity entry points)
  oxr_bvh_info bvh{};
  uint32_t spill_cap = 4;
  size_t bytes = 0;
  bool valid = false;
  float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};  // finite-vertex bounds

  ~SceneBuf() { release(); }
  void release() {
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
    dfree(d_chains);
    dfree(d_levels);
    dfree(d_texels);
    dfree(d_binds);
    dfree(d_mattex);
    dfree(d_uv[0]);
    dfree(d_uv[1]);
    dfree(d_tan);
    dfree(d_col);
    dfree(d_attr);
    dfree(d_self);
    chain_of.clear();
    valid = false;
    bytes = 0;
  }
  int upload(const oxr_scene_desc* d);
  int upload_host(const HostScene& hs);
};

int prepare_scene(const oxr_scene_desc* d, bool fast, HostScene& hs) {
  if (!d) return fail(OXR_ERR_INVALID, "null scene");
  if (d->num_tris > 0 && (!d->positions || !d->material_ids))
    return fail(OXR_ERR_INVALID, "positions and material_ids are required");
  if (d->num_tris >= (1u << 31)) return fail(OXR_ERR_INVALID, "too many triangles (%u)", d->num_tris);
  if (d->num_materials > 0 && !d->materials) return fail(OXR_ERR_INVALID, "null materials");
  if (d->num_lights > 0 && !d->lights) return fail(OXR_ERR_INVALID, "null lights");
  const uint32_t n = d->num_tris;
  {
    float l[3] = {INF_F, INF_F, INF_F}, h[3] = {-INF_F, -INF_F, -INF_F};
    for (size_t k = 0; k < (size_t)n * 9; ++k) {
      float v = d->positions[k];
      if (!std::isfinite(v)) continue;
      l[k % 3] = std::min(l[k % 3], v);
      h[k % 3] = std::max(h[k % 3], v);
    }
    for (int a = 0; a < 3; ++a) {
      hs.lo[a] = l[a] <= h[a] ? l[a] : 0.0f;
      hs.hi[a] = l[a] <= h[a] ? h[a] : 0.0f;
    }
  }
  // Materials (+ the fallback for out-of-range ids).
  std::vector<DevMaterial> mats;
  mats.reserve(d->num_materials + 1);
  for (uint32_t i = 0; i < d->num_materials; ++i) mats.push_back(prepare_material(d->materials[i]));
  const uint32_t fallback = (uint32_t)mats.size();
  const oxr_material fb = fallback_material();
  mats.push_back(prepare_material(fb));
  auto src_of = [&](uint32_t mid) -> const oxr_material& { return mid < d->num_materials ? d->materials[mid] : fb; };
  // Texture bindings per material (the fallback material has none).
  TexBuilder tb{d};
  int rc;
  if ((rc = tb.validate())) return rc;
  std::vector<DevTexBind> binds;
  std::vector<uint2> mattex(mats.size(), make_uint2(0u, 0u));
  if (d->material_textures) {
    for (uint32_t i = 0; i < d->num_materials; ++i) {
      const oxr_material_textures& mt = d->material_textures[i];
      const oxr_texture_binding* slot[TS_COUNT] = {
          &mt.base_color, &mt.metallic_roughness, &mt.normal, &mt.emissive, &mt.specular, &mt.specular_color,
          &mt.transmission, &mt.thickness, &mt.clearcoat, &mt.clearcoat_roughness, &mt.clearcoat_normal,
          &mt.sheen_color, &mt.sheen_roughness, &mt.anisotropy, &mt.iridescence, &mt.iridescence_thickness,
          &mt.diffuse_transmission, &mt.diffuse_transmission_color};
      uint32_t first = (uint32_t)binds.size();
      for (uint32_t k = 0; k < TS_COUNT; ++k) {
        const oxr_texture_binding& b = *slot[k];
        if (b.texture < 0 || (uint32_t)b.texture >= d->num_textures) continue;
        DevTexBind db{};
        db.slot = k;
        db.chain = tb.chain((uint32_t)b.texture, slot_srgb(k));
        db.uv_set = b.uv_set;
        db.xf = b.has_transform ? 1u : 0u;
        db.m00 = 1.0f;
        db.m11 = 1.0f;
        if (db.xf) {  // KHR_texture_transform: T R S
          float c = std::cos(b.rotation), s = std::sin(b.rotation);
          db.m00 = c * b.scale[0];
          db.m01 = s * b.scale[1];
          db.m02 = b.offset[0];
          db.m10 = -s * b.scale[0];
          db.m11 = c * b.scale[1];
          db.m12 = b.offset[1];
        }
        db.aux = k == TS_NORMAL ? mt.normal_scale
                 : k == TS_CC_NORMAL ? mt.clearcoat_normal_scale
                 : k == TS_IRID_THICK ? mt.iridescence_thickness_min
                                      : 0.0f;
        binds.push_back(db);
      }
      mattex[i] = make_uint2(first, (uint32_t)binds.size() - first);
    }
  }
  // Per-triangle attribute flags: an array's bit where it is given (and
  // tri_flags, when given, sets it).
  const float* attr_src[4] = {d->uv0, d->uv1, d->tangents, d->colors};
  const uint32_t attr_bit[4] = {TA_UV0, TA_UV1, TA_TANGENTS, TA_COLORS};
  bool any_attr = false;
  for (int k = 0; k < 4; ++k) any_attr = any_attr || attr_src[k] != nullptr;
  std::vector<uint32_t> attr;
  if (any_attr) {
    attr.assign(std::max(n, 1u), 0u);
    for (uint32_t g = 0; g < n; ++g)
      for (int k = 0; k < 4; ++k)
        if (attr_src[k] && (!d->tri_flags || (d->tri_flags[g] & attr_bit[k]))) attr[g] |= attr_bit[k];
  }
  std::vector<uint32_t> info(std::max(n, 1u), 0), tflags(std::max(n, 1u), 0);
  std::vector<float> alpha(std::max(n, 1u), 1.0f);
  bool has_blend = false, has_alpha_tex = false;
  for (uint32_t g = 0; g < n; ++g) {
    uint32_t mid = d->material_ids[g];
    if (mid >= d->num_materials) mid = fallback;
    bool has_n = d->normals && (d->tri_flags ? (d->tri_flags[g] & OXR_TRI_HAS_NORMALS) != 0 : true);
    info[g] = mid | (has_n ? TI_HAS_NORMALS : 0u);
    const DevMaterial& m = mats[mid];
    const oxr_material& sm = src_of(mid);
    uint32_t f = (m.flags & MAT_DS) ? TRI_DOUBLE_SIDED : 0u;
    // Alpha that varies over the triangle: a base-colour texture it can
    // sample or vertex colours (TraceScene::alpha, src/trace.rs:492-506).
    bool varies = false;
    if (sm.alpha_mode == OXR_ALPHA_MASK || sm.alpha_mode == OXR_ALPHA_BLEND) {
      uint32_t a = any_attr ? attr[g] : 0u;
      varies = (a & TA_COLORS) != 0;
      for (uint32_t j = 0; j < mattex[mid].y && !varies; ++j) {
        const DevTexBind& b = binds[mattex[mid].x + j];
        varies = b.slot == TS_BASE && b.uv_set <= 1u && (a & (b.uv_set ? TA_UV1 : TA_UV0));
      }
    }
    alpha[g] = m.base[3];
    if (sm.alpha_mode == OXR_ALPHA_MASK) {
      if (varies) {
        f |= TRI_MASK_TEX | TRI_ALPHA_TEX;
        alpha[g] = sm.alpha_cutoff;  // the record holds the cutoff
      } else if (!(m.base[3] >= sm.alpha_cutoff)) {
        f |= TRI_MASK_REJECT;
      }
    }
    if (sm.alpha_mode == OXR_ALPHA_BLEND) {
      f |= TRI_BLEND | (varies ? TRI_ALPHA_TEX : 0u);
      has_blend = true;
    }
    has_alpha_tex = has_alpha_tex || varies;
    tflags[g] = f;
  }
  // Mean emission per material: the factor times the emissive texture's
  // 1x1 level (src/pathtrace.rs:1484-1500).
  std::vector<std::array<float, 3>> mean_e(mats.size());
  for (size_t i = 0; i < mats.size(); ++i) {
    const DevMaterial& m = mats[i];
    mean_e[i] = {m.emissive[0], m.emissive[1], m.emissive[2]};
    for (uint32_t j = 0; j < mattex[i].y; ++j) {
      const DevTexBind& b = binds[mattex[i].x + j];
      if (b.slot != TS_EMISSIVE) continue;
      float4 c = tb.top(b.chain);
      mean_e[i] = {m.emissive[0] * c.x, m.emissive[1] * c.y, m.emissive[2] * c.z};
    }
  }
  // Punctual lights (zero-power lights contribute exactly nothing and are
  // dropped). Each keeps its index in the full list: the BLEND coin of its
  // shadow rays is keyed by it (src/pathtrace.rs:2147-2166).
  std::vector<DevLight> lights;
  for (uint32_t i = 0; i < d->num_lights; ++i) {
    DevLight dl;
    if (prepare_light(d->lights[i], dl)) {
      dl.index = i;
      lights.push_back(dl);
    }
  }
  if (lights.size() > 4096) return fail(OXR_ERR_INVALID, "too many punctual lights (%zu)", lights.size());
  // EmissiveLights::build (src/pathtrace.rs:1479-1543).
  std::vector<uint32_t> ltris, lookup(std::max(n, 1u), 0xffffffffu);
  std::vector<float> lpow;
  for (uint32_t g = 0; g < n; ++g) {
    const DevMaterial& m = mats[info[g] & TI_MAT_MASK];
    const std::array<float, 3>& me = mean_e[info[g] & TI_MAT_MASK];
    float e0 = (m.flags & MAT_UNLIT) ? 0.0f : me[0];
    float e1 = (m.flags & MAT_UNLIT) ? 0.0f : me[1];
    float e2 = (m.flags & MAT_UNLIT) ? 0.0f : me[2];
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
  // Leaf cap: the SAH build makes its own leaves (rarely above 2 triangles
  // at its split cost); the linear build splits down to single triangles,
  // which traced random soups 1.4-1.7x faster than leaves of up to 4
  // (20k-2M triangles) for a 1.5x longer build, still half the SAH's.
  Bvh4 bvh4 = build_bvh4(d->positions, n, tflags.data(), alpha.data(), fast ? 1 : 4, 0, fast);
  if (bvh4.nodes.empty()) {
    Node4 root;
    for (int i = 0; i < 4; ++i) {
      root.lo_x[i] = root.lo_y[i] = root.lo_z[i] = INF_F;
      root.hi_x[i] = root.hi_y[i] = root.hi_z[i] = -INF_F;
      root.child[i] = NODE_EMPTY;
      root.count[i] = 0;
    }
    bvh4.nodes.push_back(root);
  }
  hs.n = n;
  hs.mats = std::move(mats);
  hs.tex_pool = std::move(tb.pool);
  hs.tex_levels = std::move(tb.levels);
  hs.tex_chains = std::move(tb.chains);
  hs.chain_of = std::move(tb.chain_of);
  hs.binds = std::move(binds);
  hs.mattex = std::move(mattex);
  hs.any_attr = any_attr;
  hs.attr = std::move(attr);
  for (int k = 0; k < 4; ++k) hs.attr_src[k] = n > 0 ? attr_src[k] : nullptr;
  hs.pos = d->positions;
  hs.nrm = d->normals;
  hs.info = std::move(info);
  hs.has_blend = has_blend;
  hs.has_alpha_tex = has_alpha_tex;
  hs.lights = std::move(lights);
  hs.ltris = std::move(ltris);
  hs.lookup = std::move(lookup);
  hs.pmf = std::move(pmf);
  hs.cdf = std::move(cdf);
  hs.qnodes = quantize_nodes(bvh4.nodes);
  hs.tris = std::move(bvh4.tris);
  hs.num_nodes = (uint32_t)bvh4.nodes.size();
  hs.num_leaves = bvh4.num_leaves;
  hs.max_depth = bvh4.max_depth;
  hs.sah_cost = bvh4.sah_cost;
  hs.build_ms = bvh4.build_ms;
  return OXR_OK;
}

int SceneBuf::upload(const oxr_scene_desc* d) {
  HostScene hs;
  int rc = prepare_scene(d, false, hs);
  if (rc) return rc;
  return upload_host(hs);
}

int SceneBuf::upload_host(const HostScene& hs) {
  release();
  const uint32_t n = hs.n;
  for (int a = 0; a < 3; ++a) {
    lo[a] = hs.lo[a];
    hi[a] = hs.hi[a];
  }
  int rc;
  auto up = [&](auto** dst, const auto* srcp, size_t count) -> int {
    int r = dalloc(dst, count);
    if (r) return r;
    bytes += count * sizeof(**dst);
    if (count) {
      hipError_t e = hipMemcpy(*dst, srcp, count * sizeof(**dst), hipMemcpyHostToDevice);
      if (e != hipSuccess) return fail(OXR_ERR_DEVICE, "upload: %s", hipGetErrorString(e));
    }
    return OXR_OK;
  };
  if ((rc = up(&d_nodes, hs.qnodes.data(), hs.qnodes.size()))) return rc;
  if ((rc = up(&d_tris, hs.tris.data(), hs.tris.size()))) return rc;
  if ((rc = up(&d_pos, hs.pos, 9ull * n))) return rc;
  if (hs.nrm) {
    if ((rc = up(&d_nrm, hs.nrm, 9ull * n))) return rc;
  }
  if ((rc = up(&d_info, hs.info.data(), hs.info.size()))) return rc;
  if ((rc = up(&d_mats, hs.mats.data(), hs.mats.size()))) return rc;
  if ((rc = up(&d_lights, hs.lights.data(), hs.lights.size()))) return rc;
  if ((rc = up(&d_ltris, hs.ltris.data(), hs.ltris.size()))) return rc;
  if ((rc = up(&d_lpmf, hs.pmf.data(), hs.pmf.size()))) return rc;
  if ((rc = up(&d_lcdf, hs.cdf.data(), hs.cdf.size()))) return rc;
  if ((rc = up(&d_llookup, hs.lookup.data(), hs.lookup.size()))) return rc;
  const bool has_tex = !hs.binds.empty() || hs.any_attr;
  if (!hs.tex_chains.empty()) {
    if ((rc = up(&d_chains, hs.tex_chains.data(), hs.tex_chains.size()))) return rc;
    if ((rc = up(&d_levels, hs.tex_levels.data(), hs.tex_levels.size()))) return rc;
    if ((rc = up(&d_texels, hs.tex_pool.data(), hs.tex_pool.size()))) return rc;
  }
  if (!hs.binds.empty()) {
    if ((rc = up(&d_binds, hs.binds.data(), hs.binds.size()))) return rc;
    if ((rc = up(&d_mattex, hs.mattex.data(), hs.mattex.size()))) return rc;
  }
  float** attr_dst[4] = {&d_uv[0], &d_uv[1], &d_tan, &d_col};
  for (int k = 0; k < 4; ++k)
    if (hs.attr_src[k] && n > 0 && (rc = up(attr_dst[k], hs.attr_src[k], HostScene::ATTR_W[k] * n))) return rc;
  if (hs.any_attr && (rc = up(&d_attr, hs.attr.data(), hs.attr.size()))) return rc;
  chain_of = hs.chain_of;
  sc.nodes = d_nodes;
  sc.tris = d_tris;
  sc.pos = d_pos;
  sc.nrm = d_nrm;
  sc.tri_info = d_info;
  sc.mats = d_mats;
  sc.lights = d_lights;
  sc.l_tris = d_ltris;
  sc.l_pmf = d_lpmf;
  sc.l_cdf = d_lcdf;
  sc.l_lookup = d_llookup;
  sc.num_tris = n;
  sc.num_mats = (uint32_t)hs.mats.size();
  sc.num_lights = (uint32_t)hs.lights.size();
  sc.num_emissive = (uint32_t)hs.ltris.size();
  sc.has_blend = hs.has_blend ? 1u : 0u;
  sc.has_tex = has_tex ? 1u : 0u;
  sc.has_alpha_tex = hs.has_alpha_tex ? 1u : 0u;
  // Any factor that is not exactly zero (NaN included) selects the
  // extension-lobe shading; a texture only scales a factor.
  // Iridescence alone selects the full variant (2); the other extension
  // lobes the variant without the thin-film Fresnel (1).
  bool has_ext = false, has_irid = false;
  for (const DevMaterial& m : hs.mats) {
    has_ext = has_ext || !(m.transmission == 0.0f) || !(m.clearcoat == 0.0f) || !(m.sheen[0] == 0.0f) ||
              !(m.sheen[1] == 0.0f) || !(m.sheen[2] == 0.0f) || !(m.dt == 0.0f) || !(m.aniso == 0.0f);
    has_irid = has_irid || !(m.iridescence == 0.0f);
  }
  sc.has_ext = has_irid ? 2u : (has_ext ? 1u : 0u);
  sc.tex.chains = d_chains;
  sc.tex.levels = d_levels;
  sc.tex.texels = d_texels;
  sc.tex.binds = d_binds;
  sc.tex.mat = d_mattex;
  sc.tex.uv[0] = d_uv[0];
  sc.tex.uv[1] = d_uv[1];
  sc.tex.tan = d_tan;
  sc.tex.col = d_col;
  sc.tex.attr = d_attr;
  if ((rc = up(&d_self, &sc, 1))) return rc;
  bvh.num_nodes = hs.num_nodes;
  bvh.num_leaves = hs.num_leaves;
  bvh.num_tris = (uint32_t)hs.tris.size();
  bvh.max_depth = hs.max_depth;
  bvh.sah_cost = hs.sah_cost;
  bvh.build_ms = hs.build_ms;
  // Stack bound: each level leaves at most 3 siblings behind.
  spill_cap = (uint32_t)std::max(4, 4 * (int)hs.max_depth + 8 - launch::stack_lds_entries());
  valid = true;
  return OXR_OK;
}

DevFrame make_frame(const oxr_camera& cam, const oxr_pt_options& opts, const DevScene& sc,
                    const DevEnv* env = nullptr) {
  DevFrame f;
  std::memset(&f, 0, sizeof(f));
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
  // PathTracer::refine (src/pathtrace.rs:2537-2540): the cone spread is the
  // image-plane height of one pixel row, 2 half_h / height, for both
  // projections; the ray cone at the first hit is spread (orthographic) or
  // spread * t (perspective).
  f.spread = 2.0f * cam.half_h / (float)f.height;
  f.win_x0 = 0;
  f.win_y0 = 0;
  f.win_w = f.width;
  if (env && env->img) {
    f.env = *env;
    f.p_env = sc.num_emissive > 0 ? 0.5f : 1.0f;  // Integrator::p_env (src/pathtrace.rs:1915-1921)
  }
  return f;
}

// EnvironmentMap::new (src/pathtrace.rs:1575-1628): luminance x sin(theta)
// weights in f64, per-row conditional CDFs and the marginal CDF stored f32
// (last = 1), normalised texel mass; an image with no usable radiance
// samples uniformly in sin(theta) and radiates nothing.
struct EnvBuf {
  float4* d_img = nullptr;
  float* d_marginal = nullptr;
  float* d_conditional = nullptr;
  float* d_mass = nullptr;
  DevEnv env{};

  ~EnvBuf() { release(); }
  void release() {
    dfree(d_img);
    dfree(d_marginal);
    dfree(d_conditional);
    dfree(d_mass);
    env = DevEnv{};
  }
  int upload(const float* rgba, uint32_t w_in, uint32_t h_in, float intensity) {
    release();
    const uint32_t w = std::max(w_in, 1u), h = std::max(h_in, 1u);
    const bool valid = rgba && w_in > 0 && h_in > 0;
    const float PI = 3.14159265358979323846f;
    std::vector<double> wt((size_t)w * h, 0.0);
    std::vector<float4> img((size_t)w * h, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    if (valid) {
      for (size_t i = 0; i < img.size(); ++i) img[i] = make_float4(rgba[4 * i], rgba[4 * i + 1], rgba[4 * i + 2], 0.0f);
      for (uint32_t y = 0; y < h; ++y) {
        double st = (double)std::sin(PI * ((float)y + 0.5f) / (float)h);
        for (uint32_t x = 0; x < w; ++x) {
          const float4& c = img[(size_t)y * w + x];
          float l = 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
          wt[(size_t)y * w + x] = (std::isfinite(l) && l > 0.0f) ? (double)l * st : 0.0;
        }
      }
    }
    double total = 0.0;
    for (double v : wt) total += v;
    const bool uniform = !(total > 0.0 && std::isfinite(total));
    if (uniform) {
      for (size_t i = 0; i < wt.size(); ++i) wt[i] = (double)std::sin(PI * ((float)(i / w) + 0.5f) / (float)h);
      total = 0.0;
      for (double v : wt) total += v;
    }
    std::vector<float> marginal(h + 1), conditional((size_t)h * (w + 1)), mass(wt.size());
    marginal[0] = 0.0f;
    double acc = 0.0;
    for (uint32_t y = 0; y < h; ++y) {
      double row = 0.0;
      for (uint32_t x = 0; x < w; ++x) row += wt[(size_t)y * w + x];
      double racc = 0.0;
      float* c = &conditional[(size_t)y * (w + 1)];
      c[0] = 0.0f;
      for (uint32_t x = 0; x < w; ++x) {
        racc += wt[(size_t)y * w + x];
        c[x + 1] = row > 0.0 ? (float)(racc / row) : (float)(x + 1) / (float)w;
      }
      acc += row;
      marginal[y + 1] = (float)(acc / total);
    }
    marginal[h] = 1.0f;
    for (size_t i = 0; i < wt.size(); ++i) mass[i] = (float)(wt[i] / total);
    int rc;
    if ((rc = dalloc(&d_img, img.size())) || (rc = dalloc(&d_marginal, marginal.size())) ||
        (rc = dalloc(&d_conditional, conditional.size())) || (rc = dalloc(&d_mass, mass.size())))
      return rc;
    HIPC(hipMemcpy(d_img, img.data(), img.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d_marginal, marginal.data(), marginal.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d_conditional, conditional.data(), conditional.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPC(hipMemcpy(d_mass, mass.data(), mass.size() * sizeof(float), hipMemcpyHostToDevice));
    env.img = d_img;
    env.marginal = d_marginal;
    env.conditional = d_conditional;
    env.mass = d_mass;
    env.w = w;
    env.h = h;
    env.intensity = std::isfinite(intensity) ? std::fmax(intensity, 0.0f) : 0.0f;
    env.uniform = uniform ? 1u : 0u;
    return OXR_OK;
  }
};

// What one batch of a guided tracer needs: the network, the shading
// kernel's view (lobes per queue entry, training records) and the
// training-sample buffer.
struct GuideBatch {
  DevGuideNet net;
  DevGuideUse use;
  float4* samp;
  uint32_t* count;
  uint32_t samp_cap;
};

// =====================================================================
// The wavefront engine: path buffers, control block, traversal spill,
// and the batch loop shared by the tracer and the server.
// =====================================================================
struct Engine {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  PathBuffers pb{};
  Ctrl* d_ctrl = nullptr;
  uint32_t* d_spill = nullptr;
  size_t spill_alloc = 0;
  uint32_t spill_cap = 4;
  int grid_closest = 0, grid_any = 0;
  int grid_closest_alpha = 0, grid_any_alpha = 0;  // the textured-alpha variants
  uint32_t path_budget = 1u << 24;
  uint32_t max_budget = 1u << 24;
  int shade_waves = 2;  // plain single-scene shading: two waves per SIMD measured fastest
  // Persistent traversal queue split per XCD (1 or 8). Eight keeps each run
  // of neighbouring rays (pixel tiles) on one XCD and its L2: +13% on the
  // material sweep, -2% on the 12.4M-triangle terrain.
  uint32_t xcd_segments = 8;
  // Idle lanes at which a traversal wave takes new rays (1..64), for the
  // closest-hit and the occlusion kernel: 52-56 and 48 measured fastest on
  // the material sweep and the terrain (+3-5% over 64).
  uint32_t refill_closest = 56, refill_any = 48;
  bool refill_any_fixed = false;  // set by OXR_REFILL_ANY: the autotuner keeps it
  void* d_tmp = nullptr;
  size_t tmp_alloc = 0;
  bool timing = false;
  bool count_trav = false;  // closest-hit launches of single-scene batches count their fetches
  oxr_pt_stats stats{};
  double guide_infer_ms = 0.0;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  hipEvent_t ev_order[2] = {nullptr, nullptr};  // orders a caller's stream against `stream`
  struct Span {
    int cat;
    hipEvent_t a, b;
  };
  std::vector<Span> spans;

  // Target of one segment's reduction into a job accumulator.
  struct Reduce {
    uint32_t path0, npx, S;
    double* accum;
    double* sumsq;
  };

  ~Engine() {
    if (device >= 0) (void)hipSetDevice(device);
    for (auto& g : graphs_cached) (void)hipGraphExecDestroy(g.second);
    free_paths();
    dfree(d_ctrl);
    dfree(d_spill);
    if (d_tmp) (void)hipFree(d_tmp);
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    for (auto e : ev_order)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int init(int dev) {
    int rc = ensure_device(dev);
    if (rc) return rc;
    device = dev;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess || hipStreamCreate(&stream) != hipSuccess)
      return fail(OXR_ERR_DEVICE, "device %d: stream creation failed", dev);
    num_cu = prop.multiProcessorCount;
    grid_closest = num_cu * launch::trace_blocks_per_cu(false);
    grid_any = num_cu * launch::trace_blocks_per_cu(true);
    grid_closest_alpha = num_cu * launch::trace_blocks_per_cu(false, true);
    grid_any_alpha = num_cu * launch::trace_blocks_per_cu(true, true);
    if (const char* b = std::getenv("OXR_PATH_BUDGET")) {
      long v = std::atol(b);
      if (v >= 4096 && v <= (1l << 26)) path_budget = (uint32_t)v;
    }
    if (const char* gph = std::getenv("OXR_HIP_GRAPHS")) graphs = std::atoi(gph) != 0;
    if (const char* x = std::getenv("OXR_XCD")) xcd_segments = std::atoi(x) >= 8 ? 8u : 1u;
    if (const char* r = std::getenv("OXR_REFILL_CLOSEST")) {
      int v = std::atoi(r);
      if (v >= 1 && v <= 64) refill_closest = (uint32_t)v;
    }
    if (const char* r = std::getenv("OXR_REFILL_ANY")) {
      int v = std::atoi(r);
      if (v >= 1 && v <= 64) {
        refill_any = (uint32_t)v;
        refill_any_fixed = true;
      }
    }
    if (const char* w = std::getenv("OXR_SHADE_WAVES")) {
      int v = std::atoi(w);
      if (v >= 1 && v <= 3) shade_waves = v;
    }
    // Keep the wavefront buffers (~400 B per path) under 1/16 of the device.
    size_t cap_by_mem = prop.totalGlobalMem / 16 / 400;
    max_budget = (uint32_t)std::min<size_t>(std::max<size_t>(cap_by_mem, 1u << 16), 1u << 26);
    if (path_budget > max_budget) path_budget = max_budget;
    return dalloc(&d_ctrl, 1);
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
    dfree(pb.pjob);
    pb.cap = 0;
    pb.nslots = 0;
  }

  size_t path_bytes() const {
    size_t c = pb.cap, cs = (size_t)pb.cap * pb.nslots;
    return c * (2 * 32 + 2 * 4 + 16 + 5 * 16 + 16 + 4) + cs * (32 + 4 + 16 + 1);
  }

  int ensure_paths(uint32_t cap, uint32_t nslots) {
    if (pb.cap >= cap && pb.nslots == nslots) return OXR_OK;
    free_paths();
    ++gen;
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
    if ((rc = dalloc(&pb.pjob, c))) return rc;
    HIPC(hipMemsetAsync(pb.vis, 0, cs, stream));
    pb.cap = cap;
    pb.nslots = nslots;
    return OXR_OK;
  }

  int ensure_spill(uint32_t cap) {
    if (cap > spill_cap) {
      spill_cap = cap;
      ++gen;
    }
    int g = std::max(grid_closest, grid_any);
    size_t need = (size_t)g * launch::trace_block_threads() * spill_cap;
    if (need <= spill_alloc && d_spill) return OXR_OK;
    dfree(d_spill);
    spill_alloc = 0;
    ++gen;
    int rc = dalloc(&d_spill, need);
    if (rc) return rc;
    spill_alloc = need;
    return OXR_OK;
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

  hipEvent_t next_event() {
    if (ev_used == ev_pool.size()) {
      hipEvent_t e;
      (void)hipEventCreate(&e);
      ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
  }
  void span_begin(hipEvent_t& a) {
    if (!timing) return;
    a = next_event();
    (void)hipEventRecord(a, stream);
  }
  void span_end(int cat, hipEvent_t a) {
    if (!timing) return;
    hipEvent_t b = next_event();
    (void)hipEventRecord(b, stream);
    spans.push_back({cat, a, b});
  }

  void begin_stats() {
    spans.clear();
    ev_used = 0;
    stats = oxr_pt_stats{};
    launch::zero_ctrl(stream, d_ctrl);
  }

  // One batch: paths of `segs` (P in total) through the whole bounce
  // loop, then per-segment reduction into the job accumulators.
  // The bounce loop of one batch is captured once as a hipGraph and
  // replayed while its inputs (buffers, scene, jobs table, sizes, guide)
  // are unchanged; a replay is one launch instead of 4-5 per bounce.
  struct GraphKey {
    uint32_t max_bounces, nslots, P, gen;
    int grid_shade, shade_waves;
    uint32_t nseg;
    const DevJob* jobs;
    DevScene sc;
    uint32_t guided;
    GuideBatch gb;
    const DevScene* alpha_sc;
    uint32_t tex;
  };
  bool graphs = true;
  uint32_t gen = 0;  // bumped when path or spill buffers move
  // Captured bounce loops, most recent first (at most GRAPH_CACHE): a
  // server alternates between a few round shapes (one per scene, mixed).
  static constexpr size_t GRAPH_CACHE = 8;
  std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs_cached;
  uint64_t graph_captures = 0, graph_replays = 0;
  // Serving rounds: the shading grid is the full persistent grid whatever
  // the round's path count, so the captured loop does not depend on it.
  bool fixed_shade_grid = false;

  void bounce(uint32_t k, uint32_t cur, const DevJob* d_jobs, uint32_t P, uint32_t nslots, int grid_shade,
              const DevScene& sc_any, bool multi, const GuideBatch* gb) {
    launch::prep(stream, d_ctrl, cur, k > 0 ? 1u : 0u);
    hipEvent_t e1{};
    span_begin(e1);
    launch::closest(stream, alpha_sc ? grid_closest_alpha : grid_closest, sc_any, d_jobs, pb.pjob, true, multi,
                    pb.qray[cur], pb.qpid[cur], pb.meta, 0,
                    &d_ctrl->qcount[cur].v, &d_ctrl->work[0].v, 0, pb.hits, d_spill, spill_cap, &d_ctrl->error,
                    xcd_segments | (refill_closest << 8), alpha_sc, count_trav ? &d_ctrl->trav[0].v : nullptr);
    span_end(0, e1);
    if (gb && k < gb->use.kmax) {  // the guide serves the first kmax bounces
      hipEvent_t eg{};
      span_begin(eg);
      glaunch::infer(stream, gb->net, 0u, &d_ctrl->qcount[cur].v, P, nullptr, pb.qray[cur], pb.hits,
                     const_cast<GuideLobes*>(gb->use.lobes), nullptr);
      span_end(4, eg);
    }
    hipEvent_t e2{};
    span_begin(e2);
    launch::shade(stream, grid_shade, shade_waves, d_jobs, nslots, pb, d_ctrl, cur, gb ? &gb->use : nullptr,
                  tex_variant, multi_jobs, ext_variant, area_variant, k);
    span_end(2, e2);
    hipEvent_t e3{};
    span_begin(e3);
    launch::anyhit(stream, alpha_sc ? grid_any_alpha : grid_any, sc_any, d_jobs, pb.pjob, true, multi, pb.sray,
                   pb.sidx, pb.meta, nslots, 0,
                   &d_ctrl->scount.v, &d_ctrl->work[8].v, pb.vis, d_spill, spill_cap, &d_ctrl->error,
                   xcd_segments | (refill_any << 8), alpha_sc, count_trav ? &d_ctrl->trav[2].v : nullptr);
    span_end(1, e3);
  }

  // Returns true when the loop ran as a graph.
  bool bounce_graph(const DevJob* d_jobs, uint32_t P, uint32_t nslots, uint32_t max_bounces, int grid_shade,
                    const DevScene& sc_any, bool multi, const GuideBatch* gb) {
    if (!graphs || timing || count_trav || max_bounces > 32) return false;
    GraphKey key;
    std::memset(&key, 0, sizeof(key));
    key.max_bounces = max_bounces;
    key.nslots = nslots;
    key.P = gb ? P : 0u;
    key.gen = gen;
    key.grid_shade = grid_shade;
    key.shade_waves = shade_waves;
    key.nseg = xcd_segments | (refill_closest << 8) | (refill_any << 16);
    key.jobs = d_jobs;
    key.sc = sc_any;
    key.guided = (gb != nullptr ? 1u : 0u) | (multi ? 2u : 0u);
    if (gb) std::memcpy(&key.gb, gb, sizeof(GuideBatch));
    key.alpha_sc = alpha_sc;
    key.tex = (tex_variant ? 1u : 0u) | (multi_jobs ? 2u : 0u) | ((uint32_t)ext_variant << 2) | (area_variant ? 16u : 0u);
    hipGraphExec_t gexec = nullptr;
    for (size_t q = 0; q < graphs_cached.size(); ++q)
      if (std::memcmp(&key, &graphs_cached[q].first, sizeof(key)) == 0) {
        gexec = graphs_cached[q].second;
        std::rotate(graphs_cached.begin(), graphs_cached.begin() + q, graphs_cached.begin() + q + 1);
        break;
      }
    if (!gexec) {
      if (graphs_cached.size() >= GRAPH_CACHE) {
        (void)hipGraphExecDestroy(graphs_cached.back().second);
        graphs_cached.pop_back();
      }
      if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
      uint32_t cur = 0;
      for (uint32_t k = 0; k <= max_bounces; ++k, cur ^= 1u)
        bounce(k, cur, d_jobs, P, nslots, grid_shade, sc_any, multi, gb);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(stream, &g);
      if (e != hipSuccess || !g) {
        (void)hipGetLastError();
        graphs = false;  // capture unsupported here: stay eager
        return false;
      }
      e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        graphs = false;
        return false;
      }
      graphs_cached.insert(graphs_cached.begin(), {key, gexec});
      graph_captures++;
      stats.graph_captures++;
    }
    if (hipGraphLaunch(gexec, stream) != hipSuccess) return false;
    graph_replays++;
    stats.graph_replays++;
    return true;
  }

  // One batch: paths of `segs` (P in total) through the whole bounce
  // loop, then per-segment reduction into the job accumulators.
  // multi: the batch mixes scenes (each path's scene from its job); else
  // every path traces sc_any.
  // The kernel variants of the current batch: the textured shading and
  // the textured-alpha traversal (whose single-scene launches read the
  // scene's device copy through alpha_sc).
  bool tex_variant = false;
  int ext_variant = 2;  // 0: no extension lobe; 1: no iridescence; 2: every lobe
  bool area_variant = true;  // emissive triangles or an environment map in the batch
  const DevScene* alpha_sc = nullptr;
  bool multi_jobs = true;  // the batch holds paths of more than one job

  void run_batch(const DevJob* d_jobs, const SegTable& segs, uint32_t P, uint32_t nslots, uint32_t max_bounces,
                 const std::vector<Reduce>& red, const DevScene& sc_any, bool multi, int& rc,
                 const GuideBatch* gb = nullptr, bool tex = false, const DevScene* asc = nullptr, int ext = 2,
                 bool area = true) {
    tex_variant = tex;
    ext_variant = ext;
    area_variant = area;
    alpha_sc = asc;
    multi_jobs = false;  // every segment of job slot 0: the single-job shading variant
    for (uint32_t q = 0; q < segs.n; ++q) multi_jobs = multi_jobs || segs.s[q].job != 0u;
    const int grid_shade =
        fixed_shade_grid ? num_cu * 8 : (int)std::min<uint64_t>((P + 255) / 256, (uint64_t)num_cu * 8);
    hipEvent_t ea{};
    span_begin(ea);
    launch::camera(stream, d_jobs, segs, P, pb, d_ctrl);
    span_end(3, ea);
    uint32_t cur = 0;
    if (bounce_graph(d_jobs, P, nslots, max_bounces, grid_shade, sc_any, multi, gb)) {
      cur = (max_bounces + 1u) & 1u;
    } else {
      for (uint32_t k = 0; k <= max_bounces; ++k) {
        bounce(k, cur, d_jobs, P, nslots, grid_shade, sc_any, multi, gb);
        cur ^= 1u;
        // Long paths: stop launching once every path has ended.
        if (k >= 8 && (k % 4) == 0) {
          uint32_t left = 0;
          if (hipMemcpyAsync(&left, &d_ctrl->qcount[cur].v, sizeof(uint32_t), hipMemcpyDeviceToHost, stream) !=
                  hipSuccess ||
              hipStreamSynchronize(stream) != hipSuccess) {
            rc = fail(OXR_ERR_DEVICE, "queue readback failed");
            return;
          }
          if (left == 0) break;
        }
      }
    }
    launch::prep(stream, d_ctrl, cur, 1u);
    hipEvent_t e4{};
    span_begin(e4);
    launch::finalize(stream, d_jobs, segs, P, nslots, pb);
    for (const Reduce& r : red) launch::reduce(stream, r.npx, r.S, pb.out + r.path0, r.accum, r.sumsq);
    if (gb && gb->use.rec) launch::guide_collect(stream, P, gb->use, pb, gb->samp, gb->count, gb->samp_cap);
    span_end(3, e4);
    stats.paths += P;
  }

  // Sync, check the device error flag, fold the counters and event times.
  int end_stats() {
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(stream));
    Ctrl c{};
    HIPC(hipMemcpy(&c, d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost));
    if (c.error) return fail(OXR_ERR_DEVICE, "traversal stack overflow");
    stats.closest_rays = c.rays_closest;
    stats.shadow_rays = c.rays_shadow;
    stats.closest_nodes = c.trav[0].v;
    stats.closest_tris = c.trav[1].v;
    stats.anyhit_nodes = c.trav[2].v;
    stats.anyhit_tris = c.trav[3].v;
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
      } else if (s.cat == 4) {
        guide_infer_ms += ms;
      } else {
        stats.ms_other += ms;
      }
    }
    return OXR_OK;
  }
};

// =====================================================================
// The learned path guide: parameters, Adam state and training scratch on
// one device (include/oxr/guide.h).
// =====================================================================
namespace {

uint32_t pcg_u32(uint32_t x) {
  uint32_t state = x * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
  return (word >> 22) ^ word;
}
float pcg_unit(uint32_t seed, uint64_t i) {
  uint32_t h = pcg_u32(seed ^ pcg_u32((uint32_t)i ^ pcg_u32((uint32_t)(i >> 32) + 0x9e3779b9u)));
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}

}  // namespace

struct oxr_guide {
  int device = 0;
  hipStream_t stream = nullptr;
  oxr_guide_config cfg{};
  DevGuideNet net{};
  uint64_t n_grid = 0, n_total = 0;
  float* d_params = nullptr;  // [grid | mlp]
  float* d_m = nullptr;
  float* d_v = nullptr;
  float* d_grad = nullptr;
  GradCopies gcopy{};  // zero between gradient evaluations
  uint4* d_frag = nullptr;
  uint32_t t = 0;  // Adam steps
  float4* d_samp = nullptr;
  uint32_t samp_cap = 0;
  uint32_t* d_count = nullptr;
  uint32_t last_collected = 0;  // samples of the last online pass held in d_samp
  float4* d_batch = nullptr;     // one strided training batch of them
  uint32_t batch_cap = 0;
  float* d_scal = nullptr;  // inv weight sum, loss
  uint16_t* d_xs = nullptr;
  uint16_t* d_ds = nullptr;
  uint32_t act_cap = 0;  // samples
  float4* d_part = nullptr;
  float* d_bpart = nullptr;
  uint32_t part_splits = 0;
  GuideLobes* d_lobes = nullptr;
  uint32_t lobes_cap = 0;
  float4* d_rec = nullptr;
  size_t rec_cap = 0;  // float4s
  oxr_guide_stats st{};
  bool box_set = false;

  ~oxr_guide() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    dfree(d_params);
    dfree(d_m);
    dfree(d_v);
    dfree(d_grad);
    dfree(gcopy.p);
    dfree(d_frag);
    dfree(d_samp);
    dfree(d_batch);
    dfree(d_count);
    dfree(d_scal);
    dfree(d_xs);
    dfree(d_ds);
    dfree(d_part);
    dfree(d_bpart);
    dfree(d_lobes);
    dfree(d_rec);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void set_box(const float lo[3], const float hi[3]) {
    for (int a = 0; a < 3; ++a) {
      float e = hi[a] - lo[a];
      net.lo[a] = lo[a];
      net.inv[a] = e > 0.0f ? 1.0f / e : 0.0f;
    }
    box_set = true;
  }

  int init(int dev, const oxr_guide_config& c) {
    int rc = ensure_device(dev);
    if (rc) return rc;
    device = dev;
    cfg = c;
    HIPC(hipStreamCreate(&stream));
    const uint32_t T = 1u << cfg.log2_table;
    n_grid = (uint64_t)OXR_GUIDE_LEVELS * T * OXR_GUIDE_FEATURES;
    n_total = n_grid + OXR_GUIDE_MLP_PARAMS;
    double b = std::exp((std::log((double)cfg.max_res) - std::log((double)cfg.base_res)) / (OXR_GUIDE_LEVELS - 1));
    net.dense_mask = 0;
    for (int l = 0; l < OXR_GUIDE_LEVELS; ++l) {
      net.res[l] = (uint32_t)std::floor((double)cfg.base_res * std::pow(b, l) + 1e-6);
      double v = (double)net.res[l] + 1.0;
      if (v * v * v <= (double)T) net.dense_mask |= 1u << l;
    }
    net.tmask = T - 1u;
    net.log2_table = cfg.log2_table;
    gcopy = GradCopies{};
    for (int l = 0; l < G_COPY_LEVELS && ((net.dense_mask >> l) & 1u); ++l) {
      const uint64_t r = net.res[l] + 1ull;
      if (r * r * r > G_COPY_MAX_ENTRIES) break;
      gcopy.off[l] = gcopy.stride;
      gcopy.stride += (uint32_t)(2 * r * r * r);
      gcopy.nlev = (uint32_t)l + 1u;
    }
    bool empty = true;
    for (int a = 0; a < 3; ++a) empty = empty && !(cfg.aabb_max[a] > cfg.aabb_min[a]);
    if (!empty) set_box(cfg.aabb_min, cfg.aabb_max);
    if ((rc = dalloc(&d_params, n_total)) || (rc = dalloc(&d_m, n_total)) || (rc = dalloc(&d_v, n_total)) ||
        (rc = dalloc(&d_grad, n_total)) || (rc = dalloc(&d_frag, (size_t)NFRAG * 64)) || (rc = dalloc(&d_count, 4)) ||
        (rc = dalloc(&d_scal, 4)) ||
        (gcopy.nlev && (rc = dalloc(&gcopy.p, (size_t)G_GRAD_COPIES * gcopy.stride))))
      return rc;
    if (gcopy.nlev) glaunch::fill(stream, gcopy.p, (uint64_t)G_GRAD_COPIES * gcopy.stride);
    // Initialisation: grid U(-1e-4, 1e-4); hidden weights Glorot-uniform;
    // last layer x0.1 (broad, near-uniform lobes at first); biases 0.
    std::vector<float> h(n_total, 0.0f);
    for (uint64_t i = 0; i < n_grid; ++i) h[i] = (pcg_unit(cfg.seed, i) * 2.0f - 1.0f) * 1.0e-4f;
    float* mlp = h.data() + n_grid;
    const int shapes[4][3] = {{G_OW1, 64, 32}, {G_OW2, 64, 64}, {G_OW3, 64, 64}, {G_OW4, 48, 64}};
    for (int L = 0; L < 4; ++L) {
      float lim = std::sqrt(6.0f / (float)(shapes[L][1] + shapes[L][2])) * (L == 3 ? 0.1f : 1.0f);
      for (int k = 0; k < shapes[L][1] * shapes[L][2]; ++k)
        mlp[shapes[L][0] + k] = (pcg_unit(cfg.seed + 7919u * (L + 1), (uint64_t)k) * 2.0f - 1.0f) * lim;
    }
    HIPC(hipMemcpy(d_params, h.data(), n_total * sizeof(float), hipMemcpyHostToDevice));
    return reset_state();
  }

  int reset_state() {
    glaunch::fill(stream, d_m, n_total);
    glaunch::fill(stream, d_v, n_total);
    glaunch::fill(stream, d_grad, n_total);
    t = 0;
    net.grid = d_params;
    net.mlp = d_params + n_grid;
    net.frag = d_frag;
    glaunch::pack(stream, net.mlp, d_frag);
    HIPC(hipGetLastError());
    return OXR_OK;
  }

  int ensure_samples(uint32_t n) {
    if (n <= samp_cap) return OXR_OK;
    dfree(d_samp);
    samp_cap = 0;
    last_collected = 0;
    int rc = dalloc(&d_samp, 2 * (size_t)n);
    if (rc) return rc;
    samp_cap = n;
    return OXR_OK;
  }

  // Scratch of one gradient evaluation over n samples (n % 32 == 0).
  int ensure_train(uint32_t n, uint32_t nsplit) {
    int rc;
    if (n > act_cap) {
      dfree(d_xs);
      dfree(d_ds);
      act_cap = 0;
      if ((rc = dalloc(&d_xs, (size_t)(n / 16) * glaunch::xs_group())) ||
          (rc = dalloc(&d_ds, (size_t)(n / 16) * glaunch::ds_group())))
        return rc;
      act_cap = n;
    }
    if (nsplit > part_splits) {
      dfree(d_part);
      dfree(d_bpart);
      part_splits = 0;
      if ((rc = dalloc(&d_part, (size_t)nsplit * glaunch::wtiles() * 64)) ||
          (rc = dalloc(&d_bpart, (size_t)nsplit * glaunch::wtiles() * 64)))
        return rc;
      part_splits = nsplit;
    }
    return OXR_OK;
  }

  static uint32_t splits_for(uint32_t n) { return std::min<uint32_t>(64u, std::max<uint32_t>(1u, n / 32u / 32u)); }

  // Gradient of samples [samp, samp + n) into d_grad (grid part must be
  // zero on entry).
  int gradient(hipStream_t s, const float4* samp, uint32_t n) {
    uint32_t ns = splits_for(n);
    int rc = ensure_train(n, ns);
    if (rc) return rc;
    glaunch::train_grad(s, net, n, samp, d_scal, d_grad, gcopy, d_grad + n_grid, d_xs, d_ds, d_part, d_bpart, ns,
                        d_scal + 1);
    HIPC(hipGetLastError());
    return OXR_OK;
  }

  int step(hipStream_t s, const float4* samp, uint32_t n) {
    int rc = gradient(s, samp, n);
    if (rc) return rc;
    ++t;
    float c1 = 1.0f - std::pow(cfg.beta1, (float)t), c2 = 1.0f - std::pow(cfg.beta2, (float)t);
    glaunch::adam(s, n_grid, n_total, d_params, d_m, d_v, d_grad, cfg.lr, cfg.beta1, cfg.beta2, cfg.eps, c1, c2);
    glaunch::pack(s, net.mlp, d_frag);
    HIPC(hipGetLastError());
    st.steps++;
    st.samples += n;
    return OXR_OK;
  }

  int ensure_batch(uint32_t n) {
    if (n <= batch_cap) return OXR_OK;
    dfree(d_batch);
    batch_cap = 0;
    int rc = dalloc(&d_batch, 2 * (size_t)n);
    if (rc) return rc;
    batch_cap = n;
    return OXR_OK;
  }

  // Host batch -> device samples (padded with zero-weight samples to a
  // multiple of 32); returns the padded count.
  int upload_batch(uint32_t n, const float* pos, const float* dir, const float* w, uint32_t* padded) {
    uint32_t np = (n + 31u) & ~31u;
    int rc = ensure_samples(np);
    if (rc) return rc;
    last_collected = 0;  // the host batch replaces the collected samples
    std::vector<float4> h(2 * (size_t)np, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (uint32_t i = 0; i < n; ++i) {
      h[2 * i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], w ? w[i] : 0.0f);
      h[2 * i + 1] = make_float4(dir ? dir[3 * i] : 0.0f, dir ? dir[3 * i + 1] : 0.0f, dir ? dir[3 * i + 2] : 1.0f,
                                 0.0f);
    }
    for (uint32_t i = n; i < np; ++i) h[2 * i + 1] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    HIPC(hipMemcpyAsync(d_samp, h.data(), h.size() * sizeof(float4), hipMemcpyHostToDevice, stream));
    *padded = np;
    return OXR_OK;
  }

  int ensure_lobes(uint32_t n) {
    if (n <= lobes_cap) return OXR_OK;
    dfree(d_lobes);
    lobes_cap = 0;
    int rc = dalloc(&d_lobes, n);
    if (rc) return rc;
    lobes_cap = n;
    return OXR_OK;
  }

  int ensure_rec(size_t n4) {
    if (n4 <= rec_cap) return OXR_OK;
    dfree(d_rec);
    rec_cap = 0;
    int rc = dalloc(&d_rec, n4);
    if (rc) return rc;
    rec_cap = n4;
    return OXR_OK;
  }
};

// =====================================================================
// The tracer.
// =====================================================================
struct oxr_pt {
  Engine eng;
  SceneBuf scene;
  bool has_cam = false, has_opts = false;
  oxr_camera cam{};
  oxr_pt_options opts{};
  double* d_accum = nullptr;
  double* d_sumsq = nullptr;
  size_t npx_alloc = 0;
  uint32_t samples = 0;
  DevJob* d_job = nullptr;
  oxr_guide* guide = nullptr;  // not owned
  bool guide_train = false;
  EnvBuf env;

  ~oxr_pt() {
    (void)hipSetDevice(eng.device);
    (void)hipStreamSynchronize(eng.stream);
    dfree(d_accum);
    dfree(d_sumsq);
    dfree(d_job);
  }

  // Render window (x0, y0, w, h) inside the frame; w == 0: the whole frame.
  uint32_t win[4] = {0, 0, 0, 0};
  bool windowed() const { return win[2] != 0; }
  uint64_t npx_frame() const { return (uint64_t)std::max(opts.width, 1u) * std::max(opts.height, 1u); }
  // Pixels of the accumulator: the window's, or the frame's.
  uint64_t npx_accum() const { return windowed() ? (uint64_t)win[2] * win[3] : npx_frame(); }
  int check_window() const {
    if (windowed() && ((uint64_t)win[0] + win[2] > std::max(opts.width, 1u) ||
                       (uint64_t)win[1] + win[3] > std::max(opts.height, 1u)))
      return fail(OXR_ERR_INVALID, "render window %ux%u at (%u, %u) exceeds the %ux%u frame", win[2], win[3], win[0],
                  win[1], std::max(opts.width, 1u), std::max(opts.height, 1u));
    return OXR_OK;
  }

  int ensure_accum() {
    size_t npx = (size_t)npx_accum();
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
      launch::fill_f64(eng.stream, d_accum, 4 * npx_alloc);
      launch::fill_f64(eng.stream, d_sumsq, 3 * npx_alloc);
      HIPC(hipGetLastError());
    }
    return OXR_OK;
  }

  int refine(uint32_t n);
  int train_guide();
};

int oxr_pt::refine(uint32_t n) {
  if (!scene.valid || !has_cam || !has_opts) return OXR_OK;  // no-op before sync
  if (n == 0) return OXR_OK;
  n = std::min(n, std::numeric_limits<uint32_t>::max() - samples);
  if (n == 0) return OXR_OK;
  int rc;
  HIPC(hipSetDevice(eng.device));
  if ((rc = check_window())) return rc;
  if ((rc = ensure_accum())) return rc;
  const uint64_t npx = npx_accum();
  if (npx > (1ull << 31)) return fail(OXR_ERR_INVALID, "image too large (%llu pixels)", (unsigned long long)npx);
  DevJob job;
  job.sc = scene.sc;
  job.fr = make_frame(cam, opts, scene.sc, env.env.img ? &env.env : nullptr);
  if (windowed()) {
    job.fr.win_x0 = win[0];
    job.fr.win_y0 = win[1];
    job.fr.win_w = win[2];
  }
  job.accum = d_accum;
  job.sumsq = d_sumsq;
  HIPC(hipMemcpyAsync(d_job, &job, sizeof(DevJob), hipMemcpyHostToDevice, eng.stream));
  const uint32_t nslots = job.fr.nslots;
  // Batch = pixel range x sample range, at most path_budget paths (a
  // guided batch also holds 160 bytes of lobes per path: at most 4M).
  const uint32_t budget = guide ? std::min<uint32_t>(eng.path_budget, 1u << 22) : eng.path_budget;
  uint32_t npx_b = (uint32_t)std::min<uint64_t>(npx, budget);
  uint32_t S_b = npx_b < npx ? 1u : std::max(1u, std::min(n, (uint32_t)(budget / npx)));
  if ((rc = eng.ensure_paths(npx_b * S_b, nslots))) return rc;
  if ((rc = eng.ensure_spill(scene.spill_cap))) return rc;
  GuideBatch gbuf{};
  const GuideBatch* gb = nullptr;
  size_t rec_floats = 0;
  // Online training until the guide's step budget is spent; after that
  // the guide only samples (no records, no collection).
  const bool train_now =
      guide && guide_train && (guide->cfg.max_steps == 0 || guide->st.steps < guide->cfg.max_steps);
  if (guide) {
    if (!guide->box_set) guide->set_box(scene.lo, scene.hi);
    const uint32_t cap = npx_b * S_b;
    if ((rc = guide->ensure_lobes(cap))) return rc;
    if (train_now) {
      size_t nsel = ((size_t)cap + (1ull << guide->cfg.record_shift) - 1) >> guide->cfg.record_shift;
      rec_floats = nsel * G_KREC * 16;
      if ((rc = guide->ensure_rec(nsel * G_KREC * 4))) return rc;
      if ((rc = guide->ensure_samples(std::max<uint32_t>(1u << 20, guide->cfg.batch)))) return rc;
      HIPC(hipMemsetAsync(guide->d_count, 0, sizeof(uint32_t), eng.stream));
    }
    gbuf.net = guide->net;
    gbuf.use = DevGuideUse{guide->d_lobes, train_now ? guide->d_rec : nullptr, guide->cfg.mix,
                           guide->cfg.record_shift, std::min<uint32_t>(guide->cfg.bounces, (uint32_t)G_KREC)};
    gbuf.samp = guide->d_samp;
    gbuf.count = guide->d_count;
    gbuf.samp_cap = guide->samp_cap;
    gb = &gbuf;
  }
  eng.begin_stats();
  eng.guide_infer_ms = 0.0;
  for (uint32_t s0 = 0; s0 < n; s0 += S_b) {
    uint32_t S = std::min(S_b, n - s0);
    for (uint64_t px0 = 0; px0 < npx; px0 += npx_b) {
      uint32_t npx_c = (uint32_t)std::min<uint64_t>(npx_b, npx - px0);
      uint32_t P = npx_c * S;
      SegTable segs{};
      segs.n = 1;
      segs.s[0] = Seg{0u, 0u, npx_c, (uint32_t)px0, samples + s0, S};
      std::vector<Engine::Reduce> red{{0u, npx_c, S, d_accum + 4 * px0, d_sumsq + 3 * px0}};
      rc = OXR_OK;
      if (gb && gb->use.rec) glaunch::fill(eng.stream, reinterpret_cast<float*>(guide->d_rec), rec_floats);
      eng.run_batch(d_job, segs, P, nslots, job.fr.max_bounces, red, scene.sc, false, rc, gb, scene.sc.has_tex != 0,
                    scene.sc.has_alpha_tex ? scene.d_self : nullptr, (int)scene.sc.has_ext,
                    scene.sc.num_emissive > 0 || job.fr.env.img != nullptr);
      if (rc) return rc;
    }
  }
  if ((rc = eng.end_stats())) return rc;
  samples += n;
  if (guide) {
    guide->st.infer_ms += eng.guide_infer_ms;
    if (train_now && (rc = train_guide())) return rc;
  }
  return OXR_OK;
}

// Online training after a refine pass: cfg.steps Adam steps, each over a
// strided subset of all the samples the pass collected. The collection
// order follows the paths (image rows, then bounces), so a contiguous
// window would train each step on one band of the image.
int oxr_pt::train_guide() {
  uint32_t count = 0;
  HIPC(hipMemcpy(&count, guide->d_count, sizeof(uint32_t), hipMemcpyDeviceToHost));
  guide->st.collected += count;
  guide->last_collected = std::min(count, guide->samp_cap);
  uint32_t n = std::min(count, guide->samp_cap) & ~31u;
  if (n < 32) return OXR_OK;
  uint32_t B = std::min<uint32_t>(n, std::max<uint32_t>(32u, guide->cfg.batch & ~31u));
  int rc = guide->ensure_batch(B);
  if (rc) return rc;
  const uint32_t stride = n / B;  // start < stride keeps start + (B - 1) * stride < n
  // Training runs outside the bounce graph and ends in a synchronisation,
  // so its device time is always taken (two events per pass).
  hipEvent_t a = eng.next_event(), b = eng.next_event();
  (void)hipEventRecord(a, eng.stream);
  uint32_t steps = guide->cfg.steps;
  if (guide->cfg.max_steps) steps = (uint32_t)std::min<uint64_t>(steps, guide->cfg.max_steps - guide->st.steps);
  const uint64_t s0 = guide->st.steps;
  for (uint32_t k = 0; k < steps; ++k) {
    const float4* batch = guide->d_samp;
    if (stride > 1) {
      uint32_t start = (uint32_t)((s0 + k) % stride);
      glaunch::gather(eng.stream, guide->d_samp, start, stride, B, guide->d_batch);
      batch = guide->d_batch;
    }
    if ((rc = guide->step(eng.stream, batch, B))) return rc;
  }
  (void)hipEventRecord(b, eng.stream);
  HIPC(hipStreamSynchronize(eng.stream));
  float loss = 0.0f;
  HIPC(hipMemcpy(&loss, guide->d_scal + 1, sizeof(float), hipMemcpyDeviceToHost));
  guide->st.last_loss = loss;
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, a, b);
  guide->st.train_ms += ms;
  return OXR_OK;
}

// =====================================================================
// The server: resident scenes, FIFO admission by accumulator footprint,
// continuous batching of the active jobs' paths into shared rounds.
// =====================================================================
struct oxr_srv {
  Engine eng;
  std::vector<std::unique_ptr<SceneBuf>> scenes;
  struct Job {
    uint32_t id, scene;
    oxr_camera cam;
    oxr_pt_options opts;
    uint32_t state = OXR_JOB_QUEUED;
    uint32_t samples = 0, rounds = 0, slot = 0xffffffffu;
    double t_submit = 0, t_admit = 0, t_done = 0;
    double* d_accum = nullptr;
    double* d_sumsq = nullptr;
    size_t npx = 0;
  };
  std::vector<Job> jobs;
  std::vector<uint32_t> queue;   // queued job ids (FIFO)
  std::vector<uint32_t> active;  // active job ids
  std::vector<uint8_t> doomed;   // scenes to free once no job uses them
  uint64_t mem_budget = 0, mem_used = 0;  // mem_used: accumulators
  uint64_t scene_bytes = 0;               // resident scenes
  uint32_t nslots_all = 1;                // NEE slots of the widest resident scene
  DevJob* d_jobs = nullptr;      // device job table, MAX_SEGS slots
  std::vector<uint32_t> free_slots;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  oxr_srv_stats st{};
  uint32_t rr = 0;  // round-robin start for fairness when over budget

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }

  ~oxr_srv() {
    (void)hipSetDevice(eng.device);
    (void)hipStreamSynchronize(eng.stream);
    for (Job& j : jobs) {
      dfree(j.d_accum);
      dfree(j.d_sumsq);
    }
    dfree(d_jobs);
  }

  int add(std::unique_ptr<SceneBuf> sb, uint32_t* id) {
    scene_bytes += sb->bytes;
    nslots_all = std::max(nslots_all, sb->sc.num_lights + 1);
    scenes.push_back(std::move(sb));
    doomed.push_back(0);
    *id = (uint32_t)scenes.size() - 1;
    return OXR_OK;
  }

  bool in_use(uint32_t scene) const {
    for (uint32_t id : queue)
      if (jobs[id].scene == scene) return true;
    for (uint32_t id : active)
      if (jobs[id].scene == scene) return true;
    return false;
  }

  // Frees the scenes marked for removal that no queued or active job uses
  // (between rounds: nothing is in flight on the device then).
  void reap() {
    for (size_t i = 0; i < scenes.size(); ++i)
      if (doomed[i] && scenes[i] && !in_use((uint32_t)i)) {
        scene_bytes -= scenes[i]->bytes;
        scenes[i].reset();
      }
  }

  // Admission by HBM footprint: resident scenes, the shared wavefront
  // buffers and the accumulators of the active jobs, against the budget.
  int admit() {
    while (!queue.empty() && !free_slots.empty()) {
      Job& j = jobs[queue.front()];
      size_t need = j.npx * 7 * sizeof(double);
      if (mem_used + scene_bytes + eng.path_bytes() + need > mem_budget && !active.empty()) break;  // wait for memory
      int rc;
      if ((rc = dalloc(&j.d_accum, 4 * j.npx)) || (rc = dalloc(&j.d_sumsq, 3 * j.npx))) {
        dfree(j.d_accum);
        dfree(j.d_sumsq);
        if (active.empty()) return rc;
        break;
      }
      launch::fill_f64(eng.stream, j.d_accum, 4 * j.npx);
      launch::fill_f64(eng.stream, j.d_sumsq, 3 * j.npx);
      mem_used += need;
      j.slot = free_slots.back();
      free_slots.pop_back();
      DevJob dj;
      dj.sc = scenes[j.scene]->sc;
      dj.fr = make_frame(j.cam, j.opts, dj.sc);
      dj.accum = j.d_accum;
      dj.sumsq = j.d_sumsq;
      HIPC(hipMemcpyAsync(d_jobs + j.slot, &dj, sizeof(DevJob), hipMemcpyHostToDevice, eng.stream));
      j.state = OXR_JOB_ACTIVE;
      j.t_admit = now_ms();
      active.push_back(j.id);
      queue.erase(queue.begin());
    }
    return OXR_OK;
  }

  int step(uint32_t* n_active, uint32_t* n_done) {
    double t_start = now_ms();
    int rc;
    HIPC(hipSetDevice(eng.device));
    if ((rc = admit())) return rc;
    uint32_t done = 0;
    if (!active.empty()) {
      // Share the path budget: every active job contributes whole sample
      // passes over its pixels (at least one), round-robin from `rr`. The
      // round leads with the jobs of one scene (the next in turn), so it
      // runs the single-scene kernels; jobs of other scenes fill it only
      // when the lead scene's work leaves more than half of it empty.
      SegTable segs{};
      std::vector<Engine::Reduce> red;
      std::vector<std::pair<uint32_t, uint32_t>> took;  // (job id, samples)
      uint32_t P = 0, nslots = nslots_all, max_b = 0, spill = 4;
      size_t na = active.size();
      uint32_t budget = eng.path_budget;
      const uint32_t lead = jobs[active[rr % na]].scene;
      size_t n_lead = 0;
      for (uint32_t id : active) n_lead += jobs[id].scene == lead ? 1 : 0;
      for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1 && P >= budget / 2) break;
        for (size_t q = 0; q < na && segs.n < (uint32_t)MAX_SEGS; ++q) {
          Job& j = jobs[active[(rr + q) % na]];
          if ((j.scene == lead) != (pass == 0)) continue;
          uint32_t left = j.opts.samples_per_pixel - j.samples;
          uint32_t npx = (uint32_t)j.npx;
          size_t share = pass == 0 ? n_lead : na;
          uint32_t fair = std::max<uint32_t>(1u, (uint32_t)(budget / std::max<size_t>(share, 1) / std::max(npx, 1u)));
          uint32_t S = std::min(left, fair);
          if (P > 0 && (uint64_t)P + (uint64_t)npx * S > budget) {
            S = (uint32_t)((budget - P) / std::max(npx, 1u));
            if (S == 0) continue;
          }
          segs.s[segs.n++] = Seg{j.slot, P, npx, 0u, j.samples, S};
          red.push_back({P, npx, S, j.d_accum, j.d_sumsq});
          took.push_back({j.id, S});
          P += npx * S;
          max_b = std::max(max_b, std::min(j.opts.max_bounces, 65534u));
          spill = std::max(spill, scenes[j.scene]->spill_cap);
        }
      }
      // Paths end at their own job's bounce limit; up to 8 the loop always
      // runs 8 bounces so that rounds of different mixes share one graph.
      if (max_b <= 8u) max_b = 8u;
      rr = na ? (rr + 1) % (uint32_t)na : 0;
      if ((rc = eng.ensure_paths(std::max(P, 1u), nslots))) return rc;
      if ((rc = eng.ensure_spill(spill))) return rc;
      eng.begin_stats();
      rc = OXR_OK;
      bool multi = false, tex = false, area = false;
      int ext = 0;
      const DevScene* asc = nullptr;  // a mixed batch reads each path's scene through its job
      for (auto& [id, S_] : took) {
        multi = multi || jobs[id].scene != jobs[took[0].first].scene;
        const SceneBuf& sb = *scenes[jobs[id].scene];
        tex = tex || sb.sc.has_tex;
        ext = std::max(ext, (int)sb.sc.has_ext);
        area = area || sb.sc.num_emissive > 0;  // server jobs have no environment map
        if (sb.sc.has_alpha_tex && !asc) asc = sb.d_self;
      }
      // A mixed round reads each path's scene through its job: the scene
      // argument is unused, so it stays out of the graph key.
      eng.run_batch(d_jobs, segs, P, nslots, max_b, red, multi ? DevScene{} : scenes[jobs[took[0].first].scene]->sc,
                    multi, rc, nullptr, tex, asc, ext, area);
      if (rc) return rc;
      if ((rc = eng.end_stats())) return rc;
      st.rounds++;
      st.paths += P;
      st.graph_captures = (uint32_t)eng.graph_captures;
      st.graph_replays = (uint32_t)eng.graph_replays;
      if (!multi) st.single_scene_rounds++;
      double t = now_ms();
      for (auto& [id, S] : took) {
        Job& j = jobs[id];
        j.samples += S;
        j.rounds++;
        if (j.samples >= j.opts.samples_per_pixel) {
          j.state = OXR_JOB_DONE;
          j.t_done = t;
          active.erase(std::find(active.begin(), active.end(), id));
          free_slots.push_back(j.slot);
          ++done;
          st.jobs_done++;
        }
      }
    }
    reap();
    st.busy_ms += now_ms() - t_start;
    if (n_active) *n_active = (uint32_t)active.size();
    if (n_done) *n_done = done;
    return OXR_OK;
  }
};

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
  oxr_pt* h = new oxr_pt();
  int rc = h->eng.init(device);
  if (!rc) rc = dalloc(&h->d_job, 1);
  if (rc) {
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
  HIPC(hipSetDevice(h->eng.device));
  HIPC(hipStreamSynchronize(h->eng.stream));
  int rc = h->scene.upload(d);
  if (rc) return rc;
  return h->reset();
}

int oxr_pt_set_camera(oxr_pt* h, const oxr_camera* cam, int* reset_out) {
  if (!h || !cam) return fail(OXR_ERR_INVALID, "null argument");
  bool changed = !h->has_cam || std::memcmp(&h->cam, cam, sizeof(oxr_camera)) != 0;
  h->cam = *cam;
  h->has_cam = true;
  if (reset_out) *reset_out = changed ? 1 : 0;
  if (changed) {
    HIPC(hipSetDevice(h->eng.device));
    return h->reset();
  }
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
    HIPC(hipSetDevice(h->eng.device));
    int rc = h->ensure_accum();
    if (rc) return rc;
    return h->reset();
  }
  return OXR_OK;
}

int oxr_pt_set_window(oxr_pt* h, uint32_t x0, uint32_t y0, uint32_t w, uint32_t hgt) {
  if (!h) return fail(OXR_ERR_INVALID, "null handle");
  if ((w == 0) != (hgt == 0)) return fail(OXR_ERR_INVALID, "render window %ux%u: both sides or neither", w, hgt);
  uint32_t nw[4] = {w ? x0 : 0u, w ? y0 : 0u, w, hgt};
  if (std::memcmp(nw, h->win, sizeof(nw)) == 0) return OXR_OK;
  uint32_t old[4];
  std::memcpy(old, h->win, sizeof(old));
  std::memcpy(h->win, nw, sizeof(nw));
  int rc;
  if (h->has_opts && (rc = h->check_window())) {
    std::memcpy(h->win, old, sizeof(old));  // a rejected window leaves the tracer as it was
    return rc;
  }
  HIPC(hipSetDevice(h->eng.device));
  if ((rc = h->ensure_accum())) return rc;
  return h->reset();
}

int oxr_pt_reset(oxr_pt* h) {
  if (!h) return fail(OXR_ERR_INVALID, "null handle");
  HIPC(hipSetDevice(h->eng.device));
  return h->reset();
}

int oxr_pt_refine(oxr_pt* h, uint32_t n) {
  if (!h) return fail(OXR_ERR_INVALID, "null handle");
  return h->refine(n);
}

int oxr_pt_autotune(oxr_pt* h, uint32_t spp, oxr_tune_result* out) {
  if (!h) return fail(OXR_ERR_INVALID, "null handle");
  if (!h->scene.valid || !h->has_cam || !h->has_opts) return fail(OXR_ERR_STATE, "autotune needs a synced tracer");
  spp = std::max(spp, 1u);
  HIPC(hipSetDevice(h->eng.device));
  Engine& e = h->eng;
  const bool timing = e.timing, counting = e.count_trav;
  e.timing = false;
  e.count_trav = false;
  const uint64_t npx = h->npx_accum();
  const uint32_t budgets[] = {1u << 21, 1u << 22, 1u << 23, 1u << 24, 1u << 25};
  const int waves[] = {1, 2};
  const uint32_t segs[] = {1u, 8u};
  oxr_tune_result r{};
  double best = -1.0, worst = 1e300;
  uint32_t best_b = e.path_budget, best_x = e.xcd_segments;
  int best_w = e.shade_waves;
  // Paths per second of the engine's current configuration: warmed once
  // (buffers, graph capture), then the faster of two timed refines (one
  // alone is noisy).
  auto measure = [&](double& pps) -> int {
    int rc;
    if ((rc = h->reset()) || (rc = h->refine(spp))) return rc;
    HIPC(hipStreamSynchronize(e.stream));
    double s = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
      auto t0 = std::chrono::steady_clock::now();
      if ((rc = h->refine(spp))) return rc;
      HIPC(hipStreamSynchronize(e.stream));
      s = std::min(s, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    pps = (double)npx * spp / std::max(s, 1e-9);
    ++r.trials;
    worst = std::min(worst, pps);
    return OXR_OK;
  };
  for (uint32_t b : budgets) {
    if (b > e.max_budget) continue;
    for (int w : waves)
    for (uint32_t x : segs) {
      e.path_budget = b;
      e.shade_waves = w;
      e.xcd_segments = x;
      double pps = 0.0;
      int rc = measure(pps);
      if (rc) return rc;
      if (pps > best) {
        best = pps;
        best_b = b;
        best_w = w;
        best_x = x;
      }
    }
  }
  e.path_budget = best_b;
  e.shade_waves = best_w;
  e.xcd_segments = best_x;
  // Second stage at the chosen configuration: the occlusion kernel's refill
  // threshold. With the long occlusion rays of a deep tree (the
  // 12.4M-triangle terrain) a step ran 1.5% faster at 32 idle lanes than
  // at 48; the material sweep ran 1% slower.
  if (!e.refill_any_fixed) {
    const uint32_t keep = e.refill_any;
    uint32_t best_ra = keep;
    for (uint32_t ra : {32u}) {
      if (ra == keep) continue;
      e.refill_any = ra;
      double pps = 0.0;
      int rc = measure(pps);
      if (rc) return rc;
      if (pps > best) {
        best = pps;
        best_ra = ra;
      }
    }
    e.refill_any = best_ra;
  }
  if (e.pb.cap > 2ull * best_b) e.free_paths();  // drop oversized trial buffers
  e.timing = timing;
  e.count_trav = counting;
  int rc = h->reset();
  if (rc) return rc;
  r.path_budget = best_b;
  r.shade_waves = (uint32_t)best_w;
  r.xcd_segments = best_x;
  r.refill_closest = e.refill_closest;
  r.refill_any = e.refill_any;
  r.paths_per_s = best;
  r.worst_paths_per_s = r.trials ? worst : 0.0;
  if (out) *out = r;
  return OXR_OK;
}

int oxr_pt_set_environment(oxr_pt* h, const float* rgba, uint32_t width, uint32_t height, float intensity) {
  if (!h) return fail(OXR_ERR_INVALID, "null handle");
  HIPC(hipSetDevice(h->eng.device));
  HIPC(hipStreamSynchronize(h->eng.stream));
  if (!rgba) {
    h->env.release();
  } else {
    if (width == 0 || height == 0 || (uint64_t)width * height > (1ull << 28))
      return fail(OXR_ERR_INVALID, "environment size %ux%u", width, height);
    int rc = h->env.upload(rgba, width, height, intensity);
    if (rc) return rc;
  }
  return h->reset();
}

namespace {
int env_common(oxr_pt* h, uint32_t n, int do_sample, const float* in, float* dirs, float* pdf, float* le,
               uint8_t* ok) {
  if (!h || (n && (!in || !pdf))) return fail(OXR_ERR_INVALID, "null argument");
  if (!h->env.env.img) return fail(OXR_ERR_STATE, "no environment map installed");
  if (n == 0) return OXR_OK;
  HIPC(hipSetDevice(h->eng.device));
  const size_t nin = (size_t)n * (do_sample ? 2 : 3);
  float *d_in = nullptr, *d_dirs = nullptr, *d_pdf = nullptr, *d_le = nullptr;
  uint8_t* d_ok = nullptr;
  int rc;
  if ((rc = dalloc(&d_in, nin)) || (rc = dalloc(&d_dirs, 3 * (size_t)n)) || (rc = dalloc(&d_pdf, n)) ||
      (rc = dalloc(&d_le, 3 * (size_t)n)) || (rc = dalloc(&d_ok, n))) {
    dfree(d_in);
    dfree(d_dirs);
    dfree(d_pdf);
    dfree(d_le);
    return rc;
  }
  hipError_t e = hipMemcpy(d_in, in, nin * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch::env(h->eng.stream, h->env.env, n, do_sample, d_in, d_dirs, d_pdf, d_le, d_ok);
    e = hipStreamSynchronize(h->eng.stream);
  }
  if (e == hipSuccess) e = hipMemcpy(pdf, d_pdf, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess && le) e = hipMemcpy(le, d_le, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess && dirs && do_sample)
    e = hipMemcpy(dirs, d_dirs, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess && ok && do_sample) e = hipMemcpy(ok, d_ok, n, hipMemcpyDeviceToHost);
  dfree(d_in);
  dfree(d_dirs);
  dfree(d_pdf);
  dfree(d_le);
  dfree(d_ok);
  if (e != hipSuccess) return fail(OXR_ERR_DEVICE, "environment query: %s", hipGetErrorString(e));
  return OXR_OK;
}
}  // namespace

int oxr_pt_env_eval(oxr_pt* h, uint32_t n, const float* dirs, float* pdf, float* le) {
  return env_common(h, n, 0, dirs, nullptr, pdf, le, nullptr);
}

int oxr_pt_env_sample(oxr_pt* h, uint32_t n, const float* u, float* dirs, float* pdf, uint8_t* ok) {
  if (!dirs || !ok) return fail(OXR_ERR_INVALID, "null argument");
  return env_common(h, n, 1, u, dirs, pdf, nullptr, ok);
}

extern "C++" {
namespace {
// Host arrays in, one kernel, host array out, through temporary device
// buffers (the parity entry points).
struct DevTmp {
  std::vector<void*> ptrs;
  ~DevTmp() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  int in(T** d, const T* src, size_t n) {
    int rc = dalloc(d, std::max<size_t>(n, 1));
    if (rc) return rc;
    ptrs.push_back(*d);
    if (n && src && hipMemcpy(*d, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      return fail(OXR_ERR_DEVICE, "upload failed");
    return OXR_OK;
  }
};
}  // namespace
}

int oxr_pt_tex_sample(oxr_pt* h, uint32_t tex, int srgb, uint32_t n, const float* uv, const float* lod, float* out) {
  if (!h || (n && (!uv || !lod || !out))) return fail(OXR_ERR_INVALID, "null argument");
  if (!h->scene.valid) return fail(OXR_ERR_STATE, "no scene uploaded");
  size_t key = 2ull * tex + (srgb ? 0 : 1);
  if (key >= h->scene.chain_of.size() || h->scene.chain_of[key] < 0)
    return fail(OXR_ERR_INVALID, "texture %u is not used in that colour space by any material", tex);
  if (n == 0) return OXR_OK;
  HIPC(hipSetDevice(h->eng.device));
  DevTmp t;
  float *d_uv, *d_lod;
  float4* d_out;
  int rc;
  if ((rc = t.in(&d_uv, uv, 2ull * n)) || (rc = t.in(&d_lod, lod, n)) || (rc = t.in(&d_out, (const float4*)nullptr, n)))
    return rc;
  launch::tex_sample(h->eng.stream, h->scene.sc.tex, (uint32_t)h->scene.chain_of[key], n, d_uv, d_lod, d_out);
  HIPC(hipStreamSynchronize(h->eng.stream));
  HIPC(hipMemcpy(out, d_out, 16ull * n, hipMemcpyDeviceToHost));
  return OXR_OK;
}

int oxr_pt_material_eval(oxr_pt* h, uint32_t n, const uint32_t* g, const float* bary, const float* width,
                         const float* dir, const uint8_t* flip, float* out) {
  if (!h || (n && (!g || !bary || !width || !dir || !flip || !out))) return fail(OXR_ERR_INVALID, "null argument");
  if (!h->scene.valid) return fail(OXR_ERR_STATE, "no scene uploaded");
  for (uint32_t i = 0; i < n; ++i)
    if (g[i] >= h->scene.sc.num_tris) return fail(OXR_ERR_INVALID, "triangle %u out of range", g[i]);
  if (n == 0) return OXR_OK;
  HIPC(hipSetDevice(h->eng.device));
  DevTmp t;
  uint32_t* d_g;
  float *d_b, *d_w, *d_d, *d_out;
  uint8_t* d_f;
  int rc;
  if ((rc = t.in(&d_g, g, n)) || (rc = t.in(&d_b, bary, 3ull * n)) || (rc = t.in(&d_w, width, n)) ||
      (rc = t.in(&d_d, dir, 3ull * n)) || (rc = t.in(&d_f, flip, n)) ||
      (rc = t.in(&d_out, (const float*)nullptr, 43ull * n)))
    return rc;
  launch::material_eval(h->eng.stream, h->scene.sc, n, d_g, d_b, d_w, d_d, d_f, d_out);
  HIPC(hipStreamSynchronize(h->eng.stream));
  HIPC(hipMemcpy(out, d_out, 43ull * n * sizeof(float), hipMemcpyDeviceToHost));
  return OXR_OK;
}

int oxr_pt_samples(oxr_pt* h, uint32_t* out) {
  if (!h || !out) return fail(OXR_ERR_INVALID, "null argument");
  *out = h->samples;
  return OXR_OK;
}

int oxr_pt_read_accum(oxr_pt* h, double* accum, double* sum_sq) {
  if (!h || !accum) return fail(OXR_ERR_INVALID, "null argument");
  size_t npx = (size_t)h->npx_accum();
  if (!h->npx_alloc) {
    std::memset(accum, 0, npx * 4 * sizeof(double));
    if (sum_sq) std::memset(sum_sq, 0, npx * 3 * sizeof(double));
    return OXR_OK;
  }
  HIPC(hipSetDevice(h->eng.device));
  HIPC(hipStreamSynchronize(h->eng.stream));
  HIPC(hipMemcpy(accum, h->d_accum, h->npx_alloc * 4 * sizeof(double), hipMemcpyDeviceToHost));
  if (sum_sq) HIPC(hipMemcpy(sum_sq, h->d_sumsq, h->npx_alloc * 3 * sizeof(double), hipMemcpyDeviceToHost));
  return OXR_OK;
}

static int resolve_common(oxr_pt* h, const oxr_display* disp, bool rgba8, void* out) {
  if (!h || !disp || !out) return fail(OXR_ERR_INVALID, "null argument");
  size_t npx = (size_t)h->npx_accum();
  float bg[4];
  for (int k = 0; k < 3; ++k) bg[k] = srgb_to_linear((float)disp->background[k] / 255.0f);
  bg[3] = (float)disp->background[3] / 255.0f;
  HIPC(hipSetDevice(h->eng.device));
  int rc;
  if ((rc = h->ensure_accum())) return rc;
  void* d_out;
  size_t bytes = npx * (rgba8 ? 4 : 16);
  if ((rc = h->eng.tmp(bytes, &d_out))) return rc;
  launch::resolve(h->eng.stream, rgba8, (uint32_t)npx, h->samples, h->d_accum, bg, disp->background,
                  disp->tone_map, disp->exposure, d_out);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(out, d_out, bytes, hipMemcpyDeviceToHost, h->eng.stream));
  HIPC(hipStreamSynchronize(h->eng.stream));
  return OXR_OK;
}

int oxr_pt_resolve_hdr(oxr_pt* h, const oxr_display* disp, float* out) { return resolve_common(h, disp, false, out); }
int oxr_pt_resolve_rgba8(oxr_pt* h, const oxr_display* disp, uint8_t* out) {
  return resolve_common(h, disp, true, out);
}

int oxr_pt_set_timing(oxr_pt* h, int enable) {
  if (!h) return fail(OXR_ERR_INVALID, "null handle");
  h->eng.timing = (enable & OXR_TIMING_EVENTS) != 0;
  h->eng.count_trav = (enable & OXR_TIMING_COUNTS) != 0;
  return OXR_OK;
}

int oxr_pt_stats_get(oxr_pt* h, oxr_pt_stats* out) {
  if (!h || !out) return fail(OXR_ERR_INVALID, "null argument");
  *out = h->eng.stats;
  return OXR_OK;
}

int oxr_pt_bvh_info(oxr_pt* h, oxr_bvh_info* out) {
  if (!h || !out) return fail(OXR_ERR_INVALID, "null argument");
  if (!h->scene.valid) return fail(OXR_ERR_STATE, "no scene uploaded");
  *out = h->scene.bvh;
  return OXR_OK;
}

int oxr_pt_bvh_export(oxr_pt* h, void* nodes, void* tris) {
  if (!h || !nodes || !tris) return fail(OXR_ERR_INVALID, "null argument");
  if (!h->scene.valid) return fail(OXR_ERR_STATE, "no scene uploaded");
  HIPC(hipSetDevice(h->eng.device));
  HIPC(hipStreamSynchronize(h->eng.stream));
  HIPC(hipMemcpy(nodes, h->scene.d_nodes, (size_t)h->scene.bvh.num_nodes * sizeof(Node4Q), hipMemcpyDeviceToHost));
  if (h->scene.bvh.num_tris)
    HIPC(hipMemcpy(tris, h->scene.d_tris, (size_t)h->scene.bvh.num_tris * sizeof(TriRec), hipMemcpyDeviceToHost));
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
  if (!h->scene.valid) return fail(OXR_ERR_STATE, "no scene uploaded");
  if (n == 0) return OXR_OK;
  Engine& e = h->eng;
  HIPC(hipSetDevice(e.device));
  int rc;
  if ((rc = e.ensure_spill(h->scene.spill_cap))) return rc;
  size_t rbytes = (size_t)n * 32, obytes = occ ? (size_t)n : (size_t)n * 16;
  size_t off = (rbytes + 255) & ~(size_t)255;
  void* buf;
  if ((rc = e.tmp(off + obytes + 256, &buf))) return rc;
  float4* d_rays = (float4*)buf;
  void* d_out = (char*)buf + off;
  HIPC(hipMemcpyAsync(d_rays, rays, rbytes, hipMemcpyHostToDevice, e.stream));
  launch::zero_ctrl(e.stream, e.d_ctrl);
  if (occ)
    launch::anyhit(e.stream, e.grid_any, h->scene.sc, nullptr, nullptr, false, false, d_rays, nullptr, nullptr, 1, n, nullptr,
                   &e.d_ctrl->work[8].v, (uint8_t*)d_out, e.d_spill, e.spill_cap, &e.d_ctrl->error);
  else
    launch::closest(e.stream, e.grid_closest, h->scene.sc, nullptr, nullptr, false, false, d_rays, nullptr, nullptr, n,
                    nullptr, &e.d_ctrl->work[0].v, flags & 1u, (uint4*)d_out, e.d_spill, e.spill_cap, &e.d_ctrl->error);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(occ ? (void*)occ : (void*)hits, d_out, obytes, hipMemcpyDeviceToHost, e.stream));
  Ctrl c{};
  HIPC(hipMemcpyAsync(&c, e.d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, e.stream));
  HIPC(hipStreamSynchronize(e.stream));
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
  if (!h->scene.valid) return fail(OXR_ERR_STATE, "no scene uploaded");
  if (n == 0) return OXR_OK;
  Engine& e = h->eng;
  HIPC(hipSetDevice(e.device));
  int rc;
  if ((rc = e.ensure_spill(h->scene.spill_cap))) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : e.stream;
  // The launch shares the handle's control block and spill buffer with
  // refine: on a caller's stream it waits for the handle's stream, and the
  // handle's stream then waits for it.
  if (s != e.stream) {
    for (auto& ev : e.ev_order)
      if (!ev) HIPC(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPC(hipEventRecord(e.ev_order[0], e.stream));
    HIPC(hipStreamWaitEvent(s, e.ev_order[0], 0));
  }
  launch::zero_ctrl(s, e.d_ctrl);
  launch::closest(s, e.grid_closest, h->scene.sc, nullptr, nullptr, false, false, (const float4*)d_rays, nullptr, nullptr, n,
                  nullptr, &e.d_ctrl->work[0].v, 0, (uint4*)d_hits, e.d_spill, e.spill_cap, &e.d_ctrl->error);
  HIPC(hipGetLastError());
  if (s != e.stream) {
    HIPC(hipEventRecord(e.ev_order[1], s));
    HIPC(hipStreamWaitEvent(e.stream, e.ev_order[1], 0));
  }
  return OXR_OK;
}

int oxr_trace_closest_device_noalloc(oxr_pt* h, uint32_t n, const void* d_rays, void* d_hits, void* stream) {
  if (!h || (n && (!d_rays || !d_hits))) return fail(OXR_ERR_INVALID, "null argument");
  if (!h->scene.valid) return fail(OXR_ERR_STATE, "no scene uploaded");
  if (n == 0) return OXR_OK;
  Engine& e = h->eng;
  HIPC(hipSetDevice(e.device));
  unsigned char* base = static_cast<unsigned char*>(g_scratch[e.device]);
  if (!base) return fail(OXR_ERR_STATE, "device %d: fixed scratch not set up", e.device);
  // Layout: the control block, then one spill stack of spill_cap entries
  // per thread of the grid. The traversal is persistent, so a grid that
  // fits the scratch gives the same hits as the full one.
  const size_t ctrl_bytes = (sizeof(Ctrl) + 255) & ~(size_t)255;
  const uint32_t cap = h->scene.spill_cap;
  const size_t per_block = (size_t)launch::trace_block_threads() * cap * sizeof(uint32_t);
  const size_t fit = (NOALLOC_SCRATCH_BYTES - ctrl_bytes) / per_block;
  if (fit < 1)
    return fail(OXR_ERR_INVALID, "traversal needs %zu bytes of scratch per block (tree depth %u); the fixed scratch holds %zu",
                per_block + ctrl_bytes, h->scene.bvh.max_depth, NOALLOC_SCRATCH_BYTES);
  const int grid = (int)std::min<size_t>((size_t)e.grid_closest, fit);
  Ctrl* ctrl = reinterpret_cast<Ctrl*>(base);
  uint32_t* spill = reinterpret_cast<uint32_t*>(base + ctrl_bytes);
  hipStream_t s = stream ? (hipStream_t)stream : e.stream;
  launch::zero_ctrl(s, ctrl);  // a kernel, not a memset: the call may be graph-captured
  launch::closest(s, grid, h->scene.sc, nullptr, nullptr, false, false, (const float4*)d_rays, nullptr, nullptr, n,
                  nullptr, &ctrl->work[0].v, 0, (uint4*)d_hits, spill, cap, &ctrl->error);
  HIPC(hipGetLastError());
  return OXR_OK;
}

int oxr_trace_device_check(oxr_pt* h) {
  if (!h) return fail(OXR_ERR_INVALID, "null handle");
  Engine& e = h->eng;
  if (!e.d_ctrl) return OXR_OK;
  HIPC(hipSetDevice(e.device));
  Ctrl c{};
  HIPC(hipMemcpyAsync(&c, e.d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, e.stream));
  HIPC(hipStreamSynchronize(e.stream));
  if (c.error) return fail(OXR_ERR_DEVICE, "traversal stack overflow");
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

// ---- serving ----------------------------------------------------------

int oxr_srv_create(int device, uint64_t mem_budget, oxr_srv** out) {
  if (!out) return fail(OXR_ERR_INVALID, "null out");
  *out = nullptr;
  oxr_srv* s = new oxr_srv();
  int rc = s->eng.init(device);
  s->eng.fixed_shade_grid = true;
  if (!rc) rc = dalloc(&s->d_jobs, MAX_SEGS);
  if (rc) {
    delete s;
    return rc;
  }
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) tot = 1ull << 34;
  s->mem_budget = mem_budget ? mem_budget : tot / 4;
  for (int i = MAX_SEGS - 1; i >= 0; --i) s->free_slots.push_back((uint32_t)i);
  *out = s;
  return OXR_OK;
}

int oxr_srv_destroy(oxr_srv* s) {
  delete s;
  return OXR_OK;
}

int oxr_srv_add_scene(oxr_srv* s, const oxr_scene_desc* scene, uint32_t* scene_id) {
  if (!s || !scene || !scene_id) return fail(OXR_ERR_INVALID, "null argument");
  HIPC(hipSetDevice(s->eng.device));
  auto sb = std::make_unique<SceneBuf>();
  int rc = sb->upload(scene);
  if (rc) return rc;
  return s->add(std::move(sb), scene_id);
}

struct oxr_prebuilt {
  HostScene hs;
};

int oxr_scene_prebuild(const oxr_scene_desc* scene, uint32_t flags, oxr_prebuilt** out) {
  if (!scene || !out) return fail(OXR_ERR_INVALID, "null argument");
  *out = nullptr;
  auto p = std::make_unique<oxr_prebuilt>();
  int rc = prepare_scene(scene, (flags & OXR_BUILD_FAST) != 0, p->hs);
  if (rc) return rc;
  p->hs.own();
  *out = p.release();
  return OXR_OK;
}

int oxr_prebuilt_destroy(oxr_prebuilt* p) {
  delete p;
  return OXR_OK;
}

int oxr_srv_add_prebuilt(oxr_srv* s, const oxr_prebuilt* p, uint32_t* scene_id) {
  if (!s || !p || !scene_id) return fail(OXR_ERR_INVALID, "null argument");
  HIPC(hipSetDevice(s->eng.device));
  auto sb = std::make_unique<SceneBuf>();
  int rc = sb->upload_host(p->hs);
  if (rc) return rc;
  return s->add(std::move(sb), scene_id);
}

int oxr_srv_remove_scene(oxr_srv* s, uint32_t scene_id) {
  if (!s) return fail(OXR_ERR_INVALID, "null server");
  if (scene_id >= s->scenes.size() || !s->scenes[scene_id] || s->doomed[scene_id])
    return fail(OXR_ERR_INVALID, "unknown scene %u", scene_id);
  s->doomed[scene_id] = 1;
  HIPC(hipSetDevice(s->eng.device));
  HIPC(hipStreamSynchronize(s->eng.stream));
  s->reap();
  return OXR_OK;
}

int oxr_srv_submit(oxr_srv* s, uint32_t scene_id, const oxr_camera* cam, const oxr_pt_options* opts,
                   uint32_t* job_id) {
  if (!s || !cam || !opts || !job_id) return fail(OXR_ERR_INVALID, "null argument");
  if (scene_id >= s->scenes.size() || !s->scenes[scene_id] || s->doomed[scene_id])
    return fail(OXR_ERR_INVALID, "unknown scene %u", scene_id);
  if (opts->samples_per_pixel == 0) return fail(OXR_ERR_INVALID, "samples_per_pixel must be >= 1");
  uint64_t npx = (uint64_t)std::max(opts->width, 1u) * std::max(opts->height, 1u);
  if (npx > s->eng.path_budget) return fail(OXR_ERR_INVALID, "job larger than one round (%llu pixels)",
                                            (unsigned long long)npx);
  oxr_srv::Job j;
  j.id = (uint32_t)s->jobs.size();
  j.scene = scene_id;
  j.cam = *cam;
  j.opts = *opts;
  j.npx = (size_t)npx;
  j.t_submit = s->now_ms();
  s->jobs.push_back(j);
  s->queue.push_back(j.id);
  *job_id = j.id;
  return OXR_OK;
}

int oxr_srv_step(oxr_srv* s, uint32_t* active, uint32_t* done) {
  if (!s) return fail(OXR_ERR_INVALID, "null server");
  return s->step(active, done);
}

int oxr_srv_job_status(oxr_srv* s, uint32_t job_id, oxr_job_status* out) {
  if (!s || !out) return fail(OXR_ERR_INVALID, "null argument");
  if (job_id >= s->jobs.size()) return fail(OXR_ERR_INVALID, "unknown job %u", job_id);
  const oxr_srv::Job& j = s->jobs[job_id];
  out->state = j.state;
  out->samples = j.samples;
  out->target = j.opts.samples_per_pixel;
  out->rounds = j.rounds;
  out->t_submit_ms = j.t_submit;
  out->t_admit_ms = j.t_admit;
  out->t_done_ms = j.t_done;
  return OXR_OK;
}

int oxr_srv_read_accum(oxr_srv* s, uint32_t job_id, double* accum, double* sum_sq) {
  if (!s || !accum) return fail(OXR_ERR_INVALID, "null argument");
  if (job_id >= s->jobs.size()) return fail(OXR_ERR_INVALID, "unknown job %u", job_id);
  const oxr_srv::Job& j = s->jobs[job_id];
  if (!j.d_accum) return fail(OXR_ERR_STATE, "job %u has no accumulator (queued or released)", job_id);
  HIPC(hipSetDevice(s->eng.device));
  HIPC(hipStreamSynchronize(s->eng.stream));
  HIPC(hipMemcpy(accum, j.d_accum, j.npx * 4 * sizeof(double), hipMemcpyDeviceToHost));
  if (sum_sq) HIPC(hipMemcpy(sum_sq, j.d_sumsq, j.npx * 3 * sizeof(double), hipMemcpyDeviceToHost));
  return OXR_OK;
}

int oxr_srv_release(oxr_srv* s, uint32_t job_id) {
  if (!s) return fail(OXR_ERR_INVALID, "null server");
  if (job_id >= s->jobs.size()) return fail(OXR_ERR_INVALID, "unknown job %u", job_id);
  oxr_srv::Job& j = s->jobs[job_id];
  if (j.state == OXR_JOB_ACTIVE) return fail(OXR_ERR_STATE, "job %u is still rendering", job_id);
  HIPC(hipSetDevice(s->eng.device));
  HIPC(hipStreamSynchronize(s->eng.stream));
  if (j.d_accum) s->mem_used -= j.npx * 7 * sizeof(double);
  dfree(j.d_accum);
  dfree(j.d_sumsq);
  auto q = std::find(s->queue.begin(), s->queue.end(), job_id);
  if (q != s->queue.end()) s->queue.erase(q);
  return OXR_OK;
}

int oxr_srv_stats_get(oxr_srv* s, oxr_srv_stats* out) {
  if (!s || !out) return fail(OXR_ERR_INVALID, "null argument");
  *out = s->st;
  out->active = (uint32_t)s->active.size();
  out->queued = (uint32_t)s->queue.size();
  out->bytes_resident = s->mem_used + s->eng.path_bytes() + s->scene_bytes;
  out->scenes = 0;
  for (auto& sc : s->scenes) out->scenes += sc ? 1u : 0u;
  return OXR_OK;
}

// ---------------------------------------------------------------------
// Learned path guide (include/oxr/guide.h).
// ---------------------------------------------------------------------
static_assert(G_KREC == OXR_GUIDE_MAX_BOUNCES, "one training record per guided vertex");

int oxr_guide_default_config(oxr_guide_config* c) {
  if (!c) return fail(OXR_ERR_INVALID, "null config");
  *c = oxr_guide_config{};
  c->log2_table = 19;
  c->base_res = 16;
  c->max_res = 2048;
  c->lr = 5.0e-3f;
  c->beta1 = 0.9f;
  c->beta2 = 0.99f;
  c->eps = 1.0e-15f;
  c->mix = 0.5f;
  c->batch = 65536;
  c->steps = 4;
  c->record_shift = 3;
  c->bounces = 1;
  return OXR_OK;
}

int oxr_guide_create(int device, const oxr_guide_config* c, oxr_guide** out) {
  if (!c || !out) return fail(OXR_ERR_INVALID, "null argument");
  *out = nullptr;
  if (c->log2_table < 10 || c->log2_table > 22) return fail(OXR_ERR_INVALID, "log2_table must be in 10..22");
  if (c->base_res < 2 || c->max_res < c->base_res || c->max_res > (1u << 16))
    return fail(OXR_ERR_INVALID, "need 2 <= base_res <= max_res <= 65536");
  if (!(c->mix >= 0.0f && c->mix <= 1.0f)) return fail(OXR_ERR_INVALID, "mix must be in [0, 1]");
  if (!(c->lr > 0.0f) || !std::isfinite(c->lr)) return fail(OXR_ERR_INVALID, "lr must be finite and > 0");
  if (!(c->beta1 >= 0.0f && c->beta1 < 1.0f && c->beta2 >= 0.0f && c->beta2 < 1.0f) || !(c->eps >= 0.0f))
    return fail(OXR_ERR_INVALID, "Adam betas must be in [0, 1) and eps >= 0");
  if (c->batch < 32 || c->record_shift > 16) return fail(OXR_ERR_INVALID, "batch >= 32 and record_shift <= 16");
  if (c->bounces < 1 || c->bounces > OXR_GUIDE_MAX_BOUNCES)
    return fail(OXR_ERR_INVALID, "bounces must be in 1..%d", OXR_GUIDE_MAX_BOUNCES);
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(c->aabb_min[a]) || !std::isfinite(c->aabb_max[a]))
      return fail(OXR_ERR_INVALID, "aabb must be finite");
  std::unique_ptr<oxr_guide> g(new oxr_guide());
  int rc = g->init(device, *c);
  if (rc) return rc;
  *out = g.release();
  return OXR_OK;
}

int oxr_guide_destroy(oxr_guide* g) {
  delete g;
  return OXR_OK;
}

int oxr_guide_num_params(oxr_guide* g, uint64_t* n_grid, uint64_t* n_mlp) {
  if (!g) return fail(OXR_ERR_INVALID, "null guide");
  if (n_grid) *n_grid = g->n_grid;
  if (n_mlp) *n_mlp = OXR_GUIDE_MLP_PARAMS;
  return OXR_OK;
}

int oxr_guide_get_params(oxr_guide* g, float* grid, float* mlp) {
  if (!g) return fail(OXR_ERR_INVALID, "null guide");
  HIPC(hipSetDevice(g->device));
  HIPC(hipDeviceSynchronize());
  if (grid) HIPC(hipMemcpy(grid, g->d_params, g->n_grid * sizeof(float), hipMemcpyDeviceToHost));
  if (mlp) HIPC(hipMemcpy(mlp, g->d_params + g->n_grid, OXR_GUIDE_MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
  return OXR_OK;
}

int oxr_guide_set_params(oxr_guide* g, const float* grid, const float* mlp) {
  if (!g) return fail(OXR_ERR_INVALID, "null guide");
  HIPC(hipSetDevice(g->device));
  HIPC(hipDeviceSynchronize());
  if (grid) HIPC(hipMemcpy(g->d_params, grid, g->n_grid * sizeof(float), hipMemcpyHostToDevice));
  if (mlp) HIPC(hipMemcpy(g->d_params + g->n_grid, mlp, OXR_GUIDE_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice));
  int rc = g->reset_state();
  if (rc) return rc;
  HIPC(hipStreamSynchronize(g->stream));
  return OXR_OK;
}

namespace {
// Lobes of n host positions into g->d_lobes (and raw outputs into *d_raw).
int guide_infer_host(oxr_guide* g, uint32_t n, const float* pos, float** d_raw) {
  if (!g->box_set) return fail(OXR_ERR_STATE, "guide has no position box (set aabb or attach it to a tracer)");
  int rc;
  float* d_pos = nullptr;
  if ((rc = g->ensure_lobes(n)) || (rc = dalloc(&d_pos, 3 * (size_t)n))) return rc;
  if (d_raw && (rc = dalloc(d_raw, (size_t)OXR_GUIDE_OUTPUTS * n))) {
    dfree(d_pos);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(d_pos, pos, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice, g->stream);
  if (e == hipSuccess) {
    glaunch::infer(g->stream, g->net, n, nullptr, n, d_pos, nullptr, nullptr, g->d_lobes, d_raw ? *d_raw : nullptr);
    e = hipStreamSynchronize(g->stream);
  }
  dfree(d_pos);
  if (e != hipSuccess) return fail(OXR_ERR_DEVICE, "guide inference: %s", hipGetErrorString(e));
  return OXR_OK;
}
}  // namespace

int oxr_guide_infer(oxr_guide* g, uint32_t n, const float* pos, float* raw, float* lobes) {
  if (!g || (n && !pos)) return fail(OXR_ERR_INVALID, "null argument");
  if (n == 0) return OXR_OK;
  HIPC(hipSetDevice(g->device));
  float* d_raw = nullptr;
  int rc = guide_infer_host(g, n, pos, raw ? &d_raw : nullptr);
  if (rc) return rc;
  if (raw) {
    hipError_t e = hipMemcpy(raw, d_raw, (size_t)OXR_GUIDE_OUTPUTS * n * sizeof(float), hipMemcpyDeviceToHost);
    dfree(d_raw);
    if (e != hipSuccess) return fail(OXR_ERR_DEVICE, "copy: %s", hipGetErrorString(e));
  }
  if (lobes) {
    std::vector<GuideLobes> h(n);
    HIPC(hipMemcpy(h.data(), g->d_lobes, n * sizeof(GuideLobes), hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i)
      for (int k = 0; k < OXR_GUIDE_LOBES; ++k) {
        float* o = lobes + ((size_t)i * OXR_GUIDE_LOBES + k) * 5;
        o[0] = h[i].mk[k].x;
        o[1] = h[i].mk[k].y;
        o[2] = h[i].mk[k].z;
        o[3] = h[i].mk[k].w;
        o[4] = h[i].pi[k];
      }
  }
  return OXR_OK;
}

namespace {
int guide_pdf_sample(oxr_guide* g, uint32_t n, const float* pos, const float* dir, const float* u, float* dir_out,
                     float* pdf_out) {
  HIPC(hipSetDevice(g->device));
  int rc = guide_infer_host(g, n, pos, nullptr);
  if (rc) return rc;
  float *d_in = nullptr, *d_dir = nullptr, *d_pdf = nullptr;
  if ((rc = dalloc(&d_in, 3 * (size_t)n)) || (rc = dalloc(&d_dir, 3 * (size_t)n)) || (rc = dalloc(&d_pdf, n))) {
    dfree(d_in);
    dfree(d_dir);
    return rc;
  }
  hipError_t e = hipMemcpy(d_in, u ? u : dir, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    glaunch::pdf_sample(g->stream, n, g->d_lobes, u ? nullptr : d_in, u ? d_in : nullptr, d_dir, d_pdf);
    e = hipStreamSynchronize(g->stream);
  }
  if (e == hipSuccess && dir_out) e = hipMemcpy(dir_out, d_dir, 3 * (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(pdf_out, d_pdf, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
  dfree(d_in);
  dfree(d_dir);
  dfree(d_pdf);
  if (e != hipSuccess) return fail(OXR_ERR_DEVICE, "guide pdf/sample: %s", hipGetErrorString(e));
  return OXR_OK;
}
}  // namespace

int oxr_guide_pdf(oxr_guide* g, uint32_t n, const float* pos, const float* dir, float* pdf) {
  if (!g || (n && (!pos || !dir || !pdf))) return fail(OXR_ERR_INVALID, "null argument");
  if (n == 0) return OXR_OK;
  return guide_pdf_sample(g, n, pos, dir, nullptr, nullptr, pdf);
}

int oxr_guide_sample(oxr_guide* g, uint32_t n, const float* pos, const float* u, float* dir_out, float* pdf_out) {
  if (!g || (n && (!pos || !u || !dir_out || !pdf_out))) return fail(OXR_ERR_INVALID, "null argument");
  if (n == 0) return OXR_OK;
  return guide_pdf_sample(g, n, pos, nullptr, u, dir_out, pdf_out);
}

int oxr_guide_gradient(oxr_guide* g, uint32_t n, const float* pos, const float* dir, const float* weight,
                       float* grad_grid, float* grad_mlp, double* loss) {
  if (!g || (n && (!pos || !dir || !weight))) return fail(OXR_ERR_INVALID, "null argument");
  if (!g->box_set) return fail(OXR_ERR_STATE, "guide has no position box");
  HIPC(hipSetDevice(g->device));
  HIPC(hipDeviceSynchronize());
  uint32_t np = 0;
  int rc = g->upload_batch(n, pos, dir, weight, &np);
  if (rc) return rc;
  glaunch::fill(g->stream, g->d_grad, g->n_total);
  if (np && (rc = g->gradient(g->stream, g->d_samp, np))) return rc;
  HIPC(hipStreamSynchronize(g->stream));
  if (grad_grid) HIPC(hipMemcpy(grad_grid, g->d_grad, g->n_grid * sizeof(float), hipMemcpyDeviceToHost));
  if (grad_mlp)
    HIPC(hipMemcpy(grad_mlp, g->d_grad + g->n_grid, OXR_GUIDE_MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost));
  if (loss) {
    float l = 0.0f;
    if (np) HIPC(hipMemcpy(&l, g->d_scal + 1, sizeof(float), hipMemcpyDeviceToHost));
    *loss = l;
  }
  glaunch::fill(g->stream, g->d_grad, g->n_total);  // Adam expects a zero grid gradient
  HIPC(hipStreamSynchronize(g->stream));
  return OXR_OK;
}

int oxr_guide_train(oxr_guide* g, uint32_t n, const float* pos, const float* dir, const float* weight,
                    uint32_t steps, double* loss) {
  if (!g || (n && (!pos || !dir || !weight))) return fail(OXR_ERR_INVALID, "null argument");
  if (!g->box_set) return fail(OXR_ERR_STATE, "guide has no position box");
  HIPC(hipSetDevice(g->device));
  HIPC(hipDeviceSynchronize());
  uint32_t np = 0;
  int rc = g->upload_batch(n, pos, dir, weight, &np);
  if (rc) return rc;
  for (uint32_t k = 0; k < steps && np; ++k)
    if ((rc = g->step(g->stream, g->d_samp, np))) return rc;
  HIPC(hipStreamSynchronize(g->stream));
  float l = 0.0f;
  if (np && steps) HIPC(hipMemcpy(&l, g->d_scal + 1, sizeof(float), hipMemcpyDeviceToHost));
  g->st.last_loss = l;
  if (loss) *loss = l;
  return OXR_OK;
}

int oxr_guide_stats_get(oxr_guide* g, oxr_guide_stats* out) {
  if (!g || !out) return fail(OXR_ERR_INVALID, "null argument");
  *out = g->st;
  return OXR_OK;
}

int oxr_guide_collected(oxr_guide* g, uint32_t max, float* out, uint32_t* n) {
  if (!g || !n) return fail(OXR_ERR_INVALID, "null argument");
  *n = g->last_collected;
  uint32_t m = std::min(max, g->last_collected);
  if (out && m) {
    HIPC(hipSetDevice(g->device));
    HIPC(hipStreamSynchronize(g->stream));
    HIPC(hipMemcpy(out, g->d_samp, (size_t)m * 2 * sizeof(float4), hipMemcpyDeviceToHost));
  }
  return OXR_OK;
}

int oxr_pt_set_guide(oxr_pt* h, oxr_guide* g, int train) {
  if (!h) return fail(OXR_ERR_INVALID, "null handle");
  if (g && g->device != h->eng.device) return fail(OXR_ERR_INVALID, "guide and tracer live on different devices");
  HIPC(hipSetDevice(h->eng.device));
  if (g) HIPC(hipStreamSynchronize(g->stream));
  h->guide = g;
  h->guide_train = g && train != 0;
  return h->reset();
}

}  // extern "C"
