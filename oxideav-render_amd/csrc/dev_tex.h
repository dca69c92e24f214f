// Device textures of the oxr path tracer (gfx950): linear-light mip
// chains in one float4 pool, the glTF sampler (filters, wrap modes), and
// the texture slots of TraceScene::material_oriented (src/trace.rs:
// 492-606, 640-895; PreparedTexture::sample_lod, src/texture.rs:417-512).
//
// Layout in HBM: every (texture, colour space) pair a material uses is one
// chain; a chain's levels are consecutive DevTexLevel records whose
// texels (row-major float4, row 0 at v = 0) sit at `off` in the pool.
// A material's bindings are a run of DevTexBind records tagged with the
// slot they modulate.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_math.h"

namespace oxr {

struct DevTexLevel {
  uint64_t off;  // first texel in the pool
  uint32_t w, h;
};
static_assert(sizeof(DevTexLevel) == 16, "DevTexLevel is 16 bytes");

struct DevTexChain {
  uint32_t level0;   // first DevTexLevel
  uint32_t nlevels;
  uint32_t w0, h0;   // the image's size (level 0)
  uint32_t mag_linear, min_filter, wrap_s, wrap_t;
};
static_assert(sizeof(DevTexChain) == 32, "DevTexChain is 32 bytes");

// Texture slots (the order of oxr_material_textures).
enum : uint32_t {
  TS_BASE = 0, TS_MR, TS_NORMAL, TS_EMISSIVE, TS_SPEC, TS_SPEC_COLOR, TS_TRANSMISSION, TS_THICKNESS, TS_CC,
  TS_CC_ROUGH, TS_CC_NORMAL, TS_SHEEN_COLOR, TS_SHEEN_ROUGH, TS_ANISO, TS_IRID, TS_IRID_THICK, TS_DT, TS_DT_COLOR,
  TS_COUNT
};

// One binding: the chain (in the slot's colour space), the TEXCOORD set
// and the KHR_texture_transform as u' = m00 u + m01 v + m02, v' = m10 u +
// m11 v + m12. aux: the normal-map scale (normal slots) or the iridescence
// minimum thickness.
struct DevTexBind {
  uint32_t slot;
  uint32_t chain;
  uint32_t uv_set;
  uint32_t xf;
  float m00, m01, m02, m10, m11, m12;
  float aux, pad;
};
static_assert(sizeof(DevTexBind) == 48, "DevTexBind is 48 bytes");

// Per-triangle attribute flags (oxr_scene_desc tri_flags bits).
constexpr uint32_t TA_UV0 = 2u, TA_UV1 = 4u, TA_TANGENTS = 8u, TA_COLORS = 16u;

struct DevTextures {
  const DevTexChain* chains;
  const DevTexLevel* levels;
  const float4* texels;
  const DevTexBind* binds;
  const uint2* mat;         // per material: first binding, count
  const float* uv[2];       // 6 per triangle
  const float* tan;         // 12 per triangle
  const float* col;         // 12 per triangle
  const uint32_t* attr;     // TA_* per triangle (null: none)
};

// wrap_index (src/texture.rs:280-289), with i64 as the reference.
OXR_D int64_t tex_wrap(int64_t i, int64_t n, uint32_t mode) {
  if (mode == 1u) return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);  // clamp to edge
  if (mode == 2u) {                                          // mirrored repeat
    int64_t m = i % (2 * n);
    if (m < 0) m += 2 * n;
    return m >= n ? 2 * n - 1 - m : m;
  }
  int64_t r = i % n;  // repeat (rem_euclid)
  return r < 0 ? r + n : r;
}

OXR_D F4 tex_fetch(const DevTextures& T, const DevTexLevel& L, int64_t x, int64_t y, uint32_t ws, uint32_t wt) {
  int64_t xi = tex_wrap(x, (int64_t)L.w, ws), yi = tex_wrap(y, (int64_t)L.h, wt);
  float4 t = T.texels[L.off + (uint64_t)yi * L.w + (uint64_t)xi];
  return F4{{t.x, t.y, t.z, t.w}};
}

OXR_D F4 lerp4(const F4& a, const F4& b, float t) {
  F4 o;
#pragma unroll
  for (int k = 0; k < 4; ++k) o.v[k] = a.v[k] + (b.v[k] - a.v[k]) * t;
  return o;
}

// filter_level (src/texture.rs:481-512).
OXR_D F4 tex_filter(const DevTextures& T, const DevTexLevel& L, float u, float v, bool linear, uint32_t ws,
                    uint32_t wt) {
  u = isfinite(u) ? u : 0.0f;
  v = isfinite(v) ? v : 0.0f;
  u = clampr(u, -1.0e6f, 1.0e6f);
  v = clampr(v, -1.0e6f, 1.0e6f);
  float x = u * (float)L.w, y = v * (float)L.h;
  if (!linear) return tex_fetch(T, L, (int64_t)floorf(x), (int64_t)floorf(y), ws, wt);
  float xf = x - 0.5f, yf = y - 0.5f;
  float x0 = floorf(xf), y0 = floorf(yf);
  float fx = xf - x0, fy = yf - y0;
  int64_t ix = (int64_t)x0, iy = (int64_t)y0;
  F4 a = tex_fetch(T, L, ix, iy, ws, wt), b = tex_fetch(T, L, ix + 1, iy, ws, wt);
  F4 c = tex_fetch(T, L, ix, iy + 1, ws, wt), d = tex_fetch(T, L, ix + 1, iy + 1, ws, wt);
  return lerp4(lerp4(a, b, fx), lerp4(c, d, fx), fy);
}

// PreparedTexture::sample_lod (src/texture.rs:417-447). Minification
// filters: 0 nearest, 1 linear, 2 nearest-mip-nearest, 3 linear-mip-nearest,
// 4 nearest-mip-linear, 5 linear-mip-linear; the texel filter is linear
// for 1, 3 and 5.
OXR_D F4 tex_sample_lod(const DevTextures& T, uint32_t chain, float u, float v, float lod) {
  const DevTexChain C = T.chains[chain];
  const DevTexLevel* lv = T.levels + C.level0;
  if (isnan(lod) || lod <= 0.0f) return tex_filter(T, lv[0], u, v, C.mag_linear != 0u, C.wrap_s, C.wrap_t);
  uint32_t mn = C.min_filter;
  bool lin = mn == 1u || mn == 3u || mn == 5u;
  float max_level = (float)(C.nlevels - 1u);
  if (mn <= 1u) return tex_filter(T, lv[0], u, v, lin, C.wrap_s, C.wrap_t);
  if (mn <= 3u) {
    uint32_t l = (uint32_t)clampr(roundf(lod), 0.0f, max_level);
    return tex_filter(T, lv[l], u, v, lin, C.wrap_s, C.wrap_t);
  }
  float l = clampr(lod, 0.0f, max_level);
  uint32_t l0 = (uint32_t)floorf(l);
  uint32_t l1 = l0 + 1u < C.nlevels ? l0 + 1u : C.nlevels - 1u;
  float f = l - (float)l0;
  F4 a = tex_filter(T, lv[l0], u, v, lin, C.wrap_s, C.wrap_t);
  if (f <= 0.0f || l0 == l1) return a;
  F4 b = tex_filter(T, lv[l1], u, v, lin, C.wrap_s, C.wrap_t);
  return lerp4(a, b, f);
}

OXR_D void tex_transform(const DevTexBind& b, float& u, float& v) {
  if (!b.xf) return;
  float nu = b.m00 * u + b.m01 * v + b.m02;
  float nv = b.m10 * u + b.m11 * v + b.m12;
  u = nu;
  v = nv;
}

// The ray cone at the first hit (TexLod::Cone, src/pathtrace.rs:2083-2092);
// width <= 0 selects level 0 (TexLod::Base).
struct TexCone {
  float width;
  V3 dir;
};

// TraceScene::lod (src/trace.rs:518-552).
OXR_D float tex_cone_lod(const DevTextures& T, const DevTexBind& b, const float* uvs, const float* pos9,
                         const TexCone& cone) {
  float uv[3][2];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    uv[k][0] = uvs[2 * k];
    uv[k][1] = uvs[2 * k + 1];
    tex_transform(b, uv[k][0], uv[k][1]);
  }
  const DevTexChain C = T.chains[b.chain];
  float ta = (float)C.w0 * (float)C.h0 *
             fabsf((uv[1][0] - uv[0][0]) * (uv[2][1] - uv[0][1]) - (uv[2][0] - uv[0][0]) * (uv[1][1] - uv[0][1]));
  V3 p0 = v3(pos9[0], pos9[1], pos9[2]), p1 = v3(pos9[3], pos9[4], pos9[5]), p2 = v3(pos9[6], pos9[7], pos9[8]);
  V3 cr = cross(sub(p1, p0), sub(p2, p0));
  float pa = sqrtf(dot(cr, cr));
  if (!(ta > 0.0f && pa > 0.0f && cone.width > 0.0f)) return 0.0f;
  V3 n = v3(cr.x / pa, cr.y / pa, cr.z / pa);
  float cs = fmaxr(fabsf(dot(n, cone.dir)), 1.0e-3f);
  float l = 0.5f * log2f(ta / pa) + log2f(cone.width / cs);
  return isfinite(l) ? l : 0.0f;
}

// TraceScene::sample_at (src/trace.rs:554-590): false when the triangle
// has no such TEXCOORD set.
OXR_D bool tex_tap(const DevTextures& T, const DevTexBind& b, uint32_t g, uint32_t attr, V3 bary, const float* pos9,
                   const TexCone& cone, F4& out) {
  if (b.uv_set > 1u || !(attr & (b.uv_set ? TA_UV1 : TA_UV0))) return false;
  const float* uvs = T.uv[b.uv_set] + 6ull * g;
  float lod = cone.width > 0.0f ? tex_cone_lod(T, b, uvs, pos9, cone) : 0.0f;
  float u = uvs[0] * bary.x + uvs[2] * bary.y + uvs[4] * bary.z;
  float v = uvs[1] * bary.x + uvs[3] * bary.y + uvs[5] * bary.z;
  tex_transform(b, u, v);
  out = tex_sample_lod(T, b.chain, u, v, lod);
  return true;
}

OXR_D F4 tex_interp4(const float* v12, V3 b) {
  F4 o;
#pragma unroll
  for (int k = 0; k < 4; ++k) o.v[k] = v12[k] * b.x + v12[4 + k] * b.y + v12[8 + k] * b.z;
  return o;
}

// TraceScene::tangent_frame (src/trace.rs:640-663).
OXR_D bool tex_tangent_frame(const DevTextures& T, uint32_t g, uint32_t attr, V3 b, V3 n, V3& tt, V3& bt) {
  if (!(attr & TA_TANGENTS)) return false;
  const float* t12 = T.tan + 12ull * g;
  F4 tg = tex_interp4(t12, b);
  float w = t12[3] < 0.0f ? -1.0f : 1.0f;
  V3 t3 = v3(tg.v[0], tg.v[1], tg.v[2]);
  float d = dot(n, t3);
  tt = normalise(v3(t3.x - n.x * d, t3.y - n.y * d, t3.z - n.z * d));
  if (dot(tt, tt) < 0.5f) return false;
  V3 c = cross(n, tt);
  bt = v3(c.x * w, c.y * w, c.z * w);
  return true;
}

// TraceScene::perturb (src/trace.rs:665-687).
OXR_D V3 tex_perturb(const DevTextures& T, uint32_t g, uint32_t attr, V3 b, V3 n, const F4& t, float s) {
  V3 tt, bt;
  if (!tex_tangent_frame(T, g, attr, b, n, tt, bt)) return n;
  float t0 = (t.v[0] * 2.0f - 1.0f) * s, t1 = (t.v[1] * 2.0f - 1.0f) * s, t2 = t.v[2] * 2.0f - 1.0f;
  V3 nn = normalise(v3(tt.x * t0 + bt.x * t1 + n.x * t2, tt.y * t0 + bt.y * t1 + n.y * t2,
                       tt.z * t0 + bt.z * t1 + n.z * t2));
  return dot(nn, nn) > 0.5f ? nn : n;
}

// The binding of `slot` in a material's run, or null.
OXR_D const DevTexBind* tex_find(const DevTextures& T, uint2 run, uint32_t slot) {
  for (uint32_t i = 0; i < run.y; ++i)
    if (T.binds[run.x + i].slot == slot) return T.binds + run.x + i;
  return nullptr;
}

// Integrator::emission's texture factor (src/pathtrace.rs:1951-1966): the
// emissive texture at level 0, or 1 when the triangle has none.
OXR_D V3 tex_emission_factor(const DevTextures& T, uint32_t mid, uint32_t g, V3 bary) {
  if (!T.mat) return v3(1.0f, 1.0f, 1.0f);
  uint2 run = T.mat[mid];
  if (run.y == 0u) return v3(1.0f, 1.0f, 1.0f);
  const DevTexBind* b = tex_find(T, run, TS_EMISSIVE);
  F4 t;
  uint32_t attr = T.attr ? T.attr[g] : 0u;
  if (!b || !tex_tap(T, *b, g, attr, bary, nullptr, TexCone{0.0f, v3(0, 0, 0)}, t)) return v3(1.0f, 1.0f, 1.0f);
  return v3(t.v[0], t.v[1], t.v[2]);
}

// TraceScene::alpha (src/trace.rs:492-506): the base alpha times the base
// texture's alpha at level 0 and the vertex colour's alpha.
OXR_D float tex_alpha(const DevTextures& T, float base_alpha, uint32_t mid, uint32_t g, V3 bary) {
  float a = base_alpha;
  uint32_t attr = T.attr ? T.attr[g] : 0u;
  if (T.mat) {
    uint2 run = T.mat[mid];
    const DevTexBind* b = run.y ? tex_find(T, run, TS_BASE) : nullptr;
    F4 t;
    if (b && tex_tap(T, *b, g, attr, bary, nullptr, TexCone{0.0f, v3(0, 0, 0)}, t)) a *= t.v[3];
  }
  if (attr & TA_COLORS) a *= tex_interp4(T.col + 12ull * g, bary).v[3];
  return a;
}

}  // namespace oxr
