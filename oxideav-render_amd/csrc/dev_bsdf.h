// Device BSDF of the oxr path tracer: the layered glTF metallic-roughness
// model with the KHR extension lobes (oxideav-render src/pathtrace.rs
// §4 spec at :172-266, code :512-1425). Same operation order as the
// reference so the f32 results agree to the last bits the transcendental
// functions allow.
#pragma once

#include "dev_math.h"

// Linkage of the three large BSDF/light routines (eval, sample, spherical
// triangle sampling): inlined by default; a build may make them calls.
#ifndef OXR_BSDF_CALL
#define OXR_BSDF_CALL OXR_D
#endif

namespace oxr {

extern __constant__ float c_sheen[16 * 32];

constexpr float MIN_ALPHA = 1.0e-3f;

// ---- frames / hemisphere (pathtrace.rs:468-506) -----------------------
OXR_D void basis(V3 n, V3& t, V3& b) {
  float sign = copysignf(1.0f, n.z);
  float a = -1.0f / (sign + n.z);
  float bb = n.x * n.y * a;
  t = v3(1.0f + sign * n.x * n.x * a, sign * bb, -sign * n.x);
  b = v3(bb, sign + n.y * n.y * a, -n.y);
}
OXR_D V3 to_world(V3 t, V3 b, V3 n, V3 v) {
  return v3(t.x * v.x + b.x * v.y + n.x * v.z, t.y * v.x + b.y * v.y + n.y * v.z, t.z * v.x + b.z * v.y + n.z * v.z);
}
OXR_D V3 cosine_hemisphere(float u0, float u1) {
  float a = 2.0f * u0 - 1.0f;
  float b = 2.0f * u1 - 1.0f;
  float r, phi;
  if (a == 0.0f && b == 0.0f) {
    r = 0.0f;
    phi = 0.0f;
  } else if (fabsf(a) > fabsf(b)) {
    r = a;
    phi = FRAC_PI_4_F * (b / a);
  } else {
    r = b;
    phi = FRAC_PI_2_F - FRAC_PI_4_F * (a / b);
  }
  float x = r * cosf(phi), y = r * sinf(phi);
  return v3(x, y, sqrtf(fmaxr(1.0f - x * x - y * y, 0.0f)));
}

// ---- microfacet terms (pathtrace.rs:512-611) --------------------------
OXR_D float d_ggx(float n_dot_h, float a) {
  if (n_dot_h <= 0.0f) return 0.0f;
  float a2 = a * a;
  float d = n_dot_h * n_dot_h * (a2 - 1.0f) + 1.0f;
  return a2 / (PI_F * d * d);
}
OXR_D float lambda_iso(float cos, float a) {
  float c = fmaxr(fabsf(cos), 1.0e-7f);
  float a2 = a * a;
  return (sqrtf(a2 + (1.0f - a2) * c * c) / c - 1.0f) * 0.5f;
}
OXR_D float g1_iso(float cos, float a) { return 1.0f / (1.0f + lambda_iso(cos, a)); }
OXR_D float g2_iso(float cv, float cl, float a) { return 1.0f / (1.0f + lambda_iso(cv, a) + lambda_iso(cl, a)); }
OXR_D V3 schlick(V3 f0, float cos) {
  float k = powi5(1.0f - clampr(cos, 0.0f, 1.0f));
  return v3(f0.x + (1.0f - f0.x) * k, f0.y + (1.0f - f0.y) * k, f0.z + (1.0f - f0.z) * k);
}
OXR_D float schlick1(float f0, float cos) {
  float k = powi5(1.0f - clampr(cos, 0.0f, 1.0f));
  return f0 + (1.0f - f0) * k;
}
// Heitz 2018 VNDF sample; (at, ab) = (a, a) is the isotropic form.
OXR_D V3 sample_vndf2(V3 v, float at, float ab, float u0, float u1) {
  V3 vh = normalise(v3(at * v.x, ab * v.y, v.z));
  float lensq = vh.x * vh.x + vh.y * vh.y;
  V3 t1;
  if (lensq > 0.0f) {
    float il = 1.0f / sqrtf(lensq);
    t1 = v3(-vh.y * il, vh.x * il, 0.0f);
  } else {
    t1 = v3(1.0f, 0.0f, 0.0f);
  }
  V3 t2 = cross(vh, t1);
  float r = sqrtf(u0);
  float phi = 2.0f * PI_F * u1;
  float p1 = r * cosf(phi);
  float p2 = r * sinf(phi);
  float s = 0.5f * (1.0f + vh.z);
  p2 = (1.0f - s) * sqrtf(fmaxr(1.0f - p1 * p1, 0.0f)) + s * p2;
  float pz = sqrtf(fmaxr(1.0f - p1 * p1 - p2 * p2, 0.0f));
  V3 nh = v3(p1 * t1.x + p2 * t2.x + pz * vh.x, p1 * t1.y + p2 * t2.y + pz * vh.y, p1 * t1.z + p2 * t2.z + pz * vh.z);
  return normalise(v3(at * nh.x, ab * nh.y, fmaxr(nh.z, 1.0e-6f)));
}
OXR_D float vndf_reflect_pdf(V3 n, V3 v, V3 h, float a) {
  float nv = dot(n, v);
  float nh = dot(n, h);
  if (nv <= 0.0f || nh <= 0.0f || dot(v, h) <= 0.0f) return 0.0f;
  return g1_iso(nv, a) * d_ggx(nh, a) / (4.0f * nv);
}
OXR_D V3 reflect(V3 v, V3 h) {
  float d = 2.0f * dot(v, h);
  return v3(d * h.x - v.x, d * h.y - v.y, d * h.z - v.z);
}
OXR_D bool refract(V3 v, V3 h, float eta, V3& out) {
  float ci = dot(v, h);
  float s2 = eta * eta * fmaxr(1.0f - ci * ci, 0.0f);
  if (s2 >= 1.0f) return false;
  float ct = sqrtf(1.0f - s2);
  float k = eta * ci - ct;
  out = normalise(v3(-eta * v.x + k * h.x, -eta * v.y + k * h.y, -eta * v.z + k * h.z));
  return true;
}

// ---- sheen (pathtrace.rs:617-692) --------------------------------------
OXR_D float d_charlie(float n_dot_h, float a) {
  float inv = 1.0f / fmaxr(a, 1.0e-3f);
  float sin2 = fmaxr(1.0f - n_dot_h * n_dot_h, 0.0f);
  return (2.0f + inv) * powf(sin2, inv * 0.5f) / (2.0f * PI_F);
}
OXR_D float v_neubelt(float nl, float nv) { return 1.0f / fmaxr(4.0f * (nl + nv - nl * nv), 1.0e-6f); }
OXR_D float sheen_albedo(float cos, float rough) {
  float x = clampr(clampr(cos, 0.0f, 1.0f) * 32.0f - 0.5f, 0.0f, 31.0f);
  float y = clampr(clampr(rough, 0.0f, 1.0f) * 15.0f, 0.0f, 15.0f);
  float fx0 = floorf(x), fy0 = floorf(y);
  int x0 = fx0 > 0.0f ? (int)fx0 : 0, y0 = fy0 > 0.0f ? (int)fy0 : 0;
  int x1 = min(x0 + 1, 31), y1 = min(y0 + 1, 15);
  float fx = x - (float)x0, fy = y - (float)y0;
  float a = c_sheen[y0 * 32 + x0] + (c_sheen[y0 * 32 + x1] - c_sheen[y0 * 32 + x0]) * fx;
  float b = c_sheen[y1 * 32 + x0] + (c_sheen[y1 * 32 + x1] - c_sheen[y1 * 32 + x0]) * fx;
  return a + (b - a) * fy;
}

// ---- thin-film iridescence (pathtrace.rs:1326-1417) --------------------
OXR_D V3 eval_sensitivity(float opd, V3 shift) {
  float phase = 2.0f * PI_F * opd * 1.0e-9f;
  const float val[3] = {5.4856e-13f, 4.4201e-13f, 5.2481e-13f};
  const float pos[3] = {1.6810e6f, 1.7953e6f, 2.2084e6f};
  const float var[3] = {4.3278e9f, 9.3046e9f, 6.6121e9f};
  float sh[3] = {shift.x, shift.y, shift.z};
  float xyz[3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    xyz[k] = val[k] * sqrtf(2.0f * PI_F * var[k]) * cosf(pos[k] * phase + sh[k]) * expf(-phase * phase * var[k]);
  xyz[0] += 9.7470e-14f * sqrtf(2.0f * PI_F * 4.5282e9f) * cosf(2.2399e6f * phase + sh[0]) *
            expf(-4.5282e9f * phase * phase);
#pragma unroll
  for (int k = 0; k < 3; ++k) xyz[k] /= 1.0685e-7f;
  return v3(3.2404542f * xyz[0] - 1.5371385f * xyz[1] - 0.4985314f * xyz[2],
            -0.969266f * xyz[0] + 1.8760108f * xyz[1] + 0.041556f * xyz[2],
            0.0556434f * xyz[0] - 0.2040259f * xyz[1] + 1.0572252f * xyz[2]);
}

OXR_D V3 iridescent_fresnel(float outside_ior, float film_ior, V3 base_f0, float thickness_nm, float cos1) {
  cos1 = clampr(cos1, 0.0f, 1.0f);
  float r0 = powi2((film_ior - outside_ior) / (film_ior + outside_ior));
  float r12 = r0 + (1.0f - r0) * powi5(1.0f - cos1);
  float t121 = 1.0f - r12;
  float sin2_2 = powi2(outside_ior / film_ior) * (1.0f - cos1 * cos1);
  float cos2_sq = 1.0f - sin2_2;
  if (cos2_sq < 0.0f) return v3(1.0f, 1.0f, 1.0f);
  float cos2 = sqrtf(cos2_sq);
  float bf[3] = {base_f0.x, base_f0.y, base_f0.z};
  float r23[3], phi23[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float f0 = clampr(bf[k] + 1.0e-4f, 0.0f, 0.9999f);
    float sq = sqrtf(f0);
    float base_ior = (1.0f + sq) / (1.0f - sq);
    float r1 = powi2((base_ior - film_ior) / (base_ior + film_ior));
    r23[k] = r1 + (1.0f - r1) * powi5(1.0f - cos2);
    if (base_ior < film_ior) phi23[k] = PI_F;
  }
  float opd = 2.0f * film_ior * thickness_nm * cos2;
  float phi12 = film_ior < outside_ior ? PI_F : 0.0f;
  float phi21 = PI_F - phi12;
  float phi[3] = {phi21 + phi23[0], phi21 + phi23[1], phi21 + phi23[2]};
  float out[3], r123[3], rs[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    r123[k] = clampr(r12 * r23[k], 1.0e-5f, 0.9999f);
    rs[k] = t121 * t121 * r23[k] / (1.0f - r123[k]);
    out[k] = r12 + rs[k];
  }
  float cm[3] = {rs[0] - t121, rs[1] - t121, rs[2] - t121};
  for (int m = 1; m <= 2; ++m) {
#pragma unroll
    for (int k = 0; k < 3; ++k) cm[k] *= sqrtf(r123[k]);
    V3 sm = eval_sensitivity((float)m * opd, v3((float)m * phi[0], (float)m * phi[1], (float)m * phi[2]));
    out[0] += cm[0] * 2.0f * sm.x;
    out[1] += cm[1] * 2.0f * sm.y;
    out[2] += cm[2] * 2.0f * sm.z;
  }
  return v3(clampr(out[0], 0.0f, 1.0f), clampr(out[1], 0.0f, 1.0f), clampr(out[2], 0.0f, 1.0f));
}

OXR_D float dispersion_ior(float ior, float dispersion, int c) {
  float hs = (ior - 1.0f) * 0.025f * fmaxr(dispersion, 0.0f);
  if (c == 0) return fmaxr(ior - hs, 1.0f);
  if (c == 1) return fmaxr(ior, 1.0f);
  return fmaxr(ior + hs, 1.0f);
}

// ---- material inputs at a shading point (trace.rs:201-267) ------------
struct MatSample {
  V3 base;
  float metallic, roughness;
  V3 normal;
  float ior, specular;
  V3 specular_color;
  float transmission, thickness;
  float clearcoat, clearcoat_roughness;
  V3 clearcoat_normal;
  V3 sheen_color;
  float sheen_roughness;
  float anisotropy;
  V3 anisotropy_direction;
  float iridescence, iridescence_ior, iridescence_thickness;
  float diffuse_transmission;
  V3 diffuse_transmission_color;
};

enum { LOBE_DIFFUSE = 0, LOBE_SPECULAR, LOBE_TRANSMISSION, LOBE_CLEARCOAT, LOBE_SHEEN, LOBE_DIFFUSE_T, LOBES };

// The base specular microfacet: GGX with (at, ab) in the frame (t, b, n)
// (pathtrace.rs:706-811); the isotropic code path when !aniso.
struct Microfacet {
  V3 t, b, n;
  float at, ab;
  bool aniso;

  OXR_D V3 local(V3 w) const { return v3(dot(w, t), dot(w, b), dot(w, n)); }
  OXR_D float d(V3 h) const {
    if (!aniso) return d_ggx(dot(n, h), at);
    V3 hl = local(h);
    if (hl.z <= 0.0f) return 0.0f;
    float x = hl.x / at;
    float y = hl.y / ab;
    float e = x * x + y * y + hl.z * hl.z;
    return 1.0f / (PI_F * at * ab * e * e);
  }
  OXR_D float lambda(V3 w) const {
    if (!aniso) return lambda_iso(dot(n, w), at);
    V3 wl = local(w);
    float cz = fmaxr(fabsf(wl.z), 1.0e-7f);
    float a2 = powi2(at * wl.x) + powi2(ab * wl.y);
    return (sqrtf(1.0f + a2 / (cz * cz)) - 1.0f) * 0.5f;
  }
  OXR_D float g1(V3 v) const { return 1.0f / (1.0f + lambda(v)); }
  OXR_D float g2(V3 v, V3 l) const { return 1.0f / (1.0f + lambda(v) + lambda(l)); }
  OXR_D float pdf_reflect(V3 v, V3 h) const {
    if (!aniso) return vndf_reflect_pdf(n, v, h, at);
    float nv = dot(n, v);
    float nh = dot(n, h);
    if (nv <= 0.0f || nh <= 0.0f || dot(v, h) <= 0.0f) return 0.0f;
    return g1(v) * d(h) / (4.0f * nv);
  }
  OXR_D bool sample(V3 v, float u0, float u1, V3& out) const {
    V3 vl = local(v);
    if (vl.z <= 0.0f) return false;
    V3 h = aniso ? sample_vndf2(vl, at, ab, u0, u1) : sample_vndf2(vl, at, at, u0, u1);
    out = to_world(t, b, n, h);
    return true;
  }
};

// EXT = 0: the scene has no material with clearcoat, sheen, iridescence,
// transmission, diffuse transmission or anisotropy. Those factors are then
// compile-time zeros: every branch they guard takes the way a zero factor
// takes at run time, and their lobe probabilities, which a zero factor
// makes exactly 0, are constants, so the code and the state they need drop
// out and the results stay bit for bit those of EXT = 2.
// EXT = 1: the same for iridescence alone (every material's iridescence
// factor is 0). The thin-film Fresnel is the heaviest part of the BSDF:
// with it inlined at every Fresnel evaluation the shading kernel of a
// scene with other extension lobes and emissive triangles spilled about
// 400 bytes a lane at two waves per SIMD, against about 160 without it.
// EXT = 2: every lobe.
template <int EXT>
struct BsdfT {
  V3 v, n, ng, ncc, base;
  float metallic;
  Microfacet mf;
  float specular;
  V3 f0d;
  float transmission;
  bool thin;
  float eta_i, eta_t;
  float clearcoat, cc_alpha;
  V3 sheen;
  float sheen_rough, sheen_alpha;
  float irid, irid_ior, irid_thick;
  float dt;
  V3 dt_color;
  float w_cc;
  float probs[LOBES];

  // Bsdf::new (pathtrace.rs:853-986).
  OXR_D void init(const MatSample& m, V3 v_, V3 ng_in, bool front) {
    v = v_;
    ng = front ? ng_in : neg(ng_in);
    n = front ? m.normal : neg(m.normal);
    if (dot(n, v) <= 1.0e-4f) n = ng;
    ncc = front ? m.clearcoat_normal : neg(m.clearcoat_normal);
    if (dot(ncc, v) <= 1.0e-4f) ncc = n;
    thin = m.thickness <= 0.0f;
    float ior = (isfinite(m.ior) && m.ior >= 1.0f) ? m.ior : 1.5f;
    if (thin || front) {
      eta_i = 1.0f;
      eta_t = ior;
    } else {
      eta_i = ior;
      eta_t = 1.0f;
    }
    bool inside = !thin && !front;
    float r0 = powi2((ior - 1.0f) / (ior + 1.0f));
    base = m.base;
    float alpha = fmaxr(m.roughness * m.roughness, MIN_ALPHA);
    float aniso = EXT ? clampr(m.anisotropy, 0.0f, 1.0f) : 0.0f;
    basis(n, mf.t, mf.b);
    mf.n = n;
    mf.at = alpha;
    mf.ab = alpha;
    mf.aniso = false;
    if (EXT && aniso > 0.0f) {
      V3 d0 = m.anisotropy_direction;
      float dn = dot(d0, n);
      V3 t = normalise(v3(d0.x - n.x * dn, d0.y - n.y * dn, d0.z - n.z * dn));
      if (dot(t, t) > 0.5f) {
        float at = alpha + (1.0f - alpha) * aniso * aniso;
        if (at > alpha) {
          mf.t = t;
          mf.b = cross(n, t);
          mf.at = at;
          mf.ab = alpha;
          mf.aniso = true;
        }
      }
    }
    metallic = m.metallic;
    specular = clampr(m.specular, 0.0f, 1.0f);
    f0d = v3(fminr(r0 * m.specular_color.x, 1.0f), fminr(r0 * m.specular_color.y, 1.0f),
             fminr(r0 * m.specular_color.z, 1.0f));
    if (EXT) {
      transmission = m.transmission;
      clearcoat = inside ? 0.0f : m.clearcoat;
      cc_alpha = fmaxr(m.clearcoat_roughness * m.clearcoat_roughness, MIN_ALPHA);
      sheen = inside ? v3(0.0f, 0.0f, 0.0f) : m.sheen_color;
      sheen_rough = m.sheen_roughness;
      sheen_alpha = fmaxr(m.sheen_roughness * m.sheen_roughness, 1.0e-3f);
      irid = (EXT < 2 || inside) ? 0.0f : clampr(m.iridescence, 0.0f, 1.0f);
      irid_ior = EXT < 2 ? 0.0f : (isfinite(m.iridescence_ior) ? fmaxr(m.iridescence_ior, 1.0f) : 1.3f);
      irid_thick = EXT < 2 ? 0.0f : (isfinite(m.iridescence_thickness) ? fmaxr(m.iridescence_thickness, 0.0f) : 400.0f);
      dt = clampr(m.diffuse_transmission, 0.0f, 1.0f);
      dt_color = v3(clampr(m.diffuse_transmission_color.x, 0.0f, 1.0f),
                    clampr(m.diffuse_transmission_color.y, 0.0f, 1.0f),
                    clampr(m.diffuse_transmission_color.z, 0.0f, 1.0f));
    } else {
      transmission = clearcoat = cc_alpha = sheen_rough = sheen_alpha = irid = irid_ior = irid_thick = dt = 0.0f;
      sheen = dt_color = v3(0.0f, 0.0f, 0.0f);
    }
    float cv = fmaxr(dot(n, v), 1.0e-4f);
    w_cc = EXT && clearcoat > 0.0f ? clearcoat * schlick1(0.04f, dot(ncc, v)) : 0.0f;
    float sheen_max = max3(sheen);
    float e_v = EXT && sheen_max > 0.0f ? sheen_albedo(cv, sheen_rough) : 0.0f;
    float s_scale = (1.0f - w_cc) * (1.0f - sheen_max * e_v);
    V3 fd_v, fm_v;
    fresnel_pair(cv, fd_v, fm_v);
    float fd = max3(fd_v);
    float fm = mean3(fm_v);
    float m_ = metallic;
    float base_mean = fmaxr(mean3(base), 0.0f);
    float t = transmission;
    float p0 = s_scale * (1.0f - m_) * (1.0f - t) * (1.0f - dt) * (1.0f - fd) * base_mean;
    float p1 = s_scale * ((1.0f - m_) * fd + m_ * fm);
    float p2 = s_scale * (1.0f - m_) * t * (1.0f - fd) * base_mean;
    float p3 = w_cc;
    float p4 = (1.0f - w_cc) * sheen_max * e_v;
    float p5 = s_scale * (1.0f - m_) * (1.0f - t) * dt * (1.0f - fd) * fmaxr(mean3(dt_color), 0.0f);
    float sum = p0 + p1 + p2 + p3 + p4 + p5;
    if (sum > 0.0f && isfinite(sum)) {
      probs[0] = p0 / sum;
      probs[1] = p1 / sum;
      // p2..p5 carry a zero factor (transmission, w_cc, sheen, dt) when
      // !EXT: exactly 0 over a finite positive sum.
      probs[2] = EXT ? p2 / sum : 0.0f;
      probs[3] = EXT ? p3 / sum : 0.0f;
      probs[4] = EXT ? p4 / sum : 0.0f;
      probs[5] = EXT ? p5 / sum : 0.0f;
    } else {
#pragma unroll
      for (int i = 0; i < LOBES; ++i) probs[i] = 0.0f;
    }
  }

  OXR_D bool black() const {
    return !(probs[0] > 0.0f || probs[1] > 0.0f || probs[2] > 0.0f || probs[3] > 0.0f || probs[4] > 0.0f ||
             probs[5] > 0.0f);
  }

  OXR_D V3 fresnel_d(float c) const {
    if (eta_i > eta_t) {
      float eta = eta_i / eta_t;
      float s2 = eta * eta * fmaxr(1.0f - c * c, 0.0f);
      if (s2 >= 1.0f) return v3(1.0f, 1.0f, 1.0f);
      return schlick(f0d, sqrtf(1.0f - s2));
    }
    return schlick(f0d, c);
  }

  OXR_D void fresnel_pair(float c, V3& fd_out, V3& fm_out) const {
    V3 fd = scale(fresnel_d(c), specular);
    V3 fm = schlick(base, c);
    if (EXT < 2 || irid <= 0.0f) {
      fd_out = fd;
      fm_out = fm;
      return;
    }
    float i = irid;
    V3 fd_i = metallic < 1.0f ? iridescent_fresnel(1.0f, irid_ior, scale(f0d, specular), irid_thick, c) : fd;
    V3 fm_i = metallic > 0.0f ? iridescent_fresnel(1.0f, irid_ior, base, irid_thick, c) : fm;
    fd_out = v3(fd.x + (fd_i.x - fd.x) * i, fd.y + (fd_i.y - fd.y) * i, fd.z + (fd_i.z - fd.z) * i);
    fm_out = v3(fm.x + (fm_i.x - fm.x) * i, fm.y + (fm_i.y - fm.y) * i, fm.z + (fm_i.z - fm.z) * i);
  }

  // Bsdf::eval (pathtrace.rs:1075-1211): f·|n·l| and the mixture pdf.
  OXR_BSDF_CALL void eval(V3 l, V3& f_out, float& pdf_out) const {
    float nv = fmaxr(dot(n, v), 1.0e-4f);
    if (dot(ng, l) > 0.0f) {
      float nl = dot(n, l);
      V3 h = normalise(add(v, l));
      float vh = fmaxr(dot(v, h), 0.0f);
      V3 f = v3(0.0f, 0.0f, 0.0f);
      float pdf = 0.0f;
      if (nl > 0.0f) {
        V3 fd, fm;
        fresnel_pair(vh, fd, fm);
        float spec_c = mf.d(h) * mf.g2(v, l) / (4.0f * nv * nl);
        float m = metallic;
        float diff_k = (1.0f - m) * (1.0f - transmission) * (1.0f - dt) * (1.0f - max3(fd)) / PI_F;
        V3 layer = v3(spec_c * ((1.0f - m) * fd.x + m * fm.x) + diff_k * base.x,
                      spec_c * ((1.0f - m) * fd.y + m * fm.y) + diff_k * base.y,
                      spec_c * ((1.0f - m) * fd.z + m * fm.z) + diff_k * base.z);
        float smax = max3(sheen);
        if (EXT && smax > 0.0f) {
          float sc = fminr(1.0f - smax * sheen_albedo(nv, sheen_rough), 1.0f - smax * sheen_albedo(nl, sheen_rough));
          float nh = dot(n, h);
          float sh = d_charlie(nh, sheen_alpha) * v_neubelt(nl, nv);
          layer = v3(layer.x * sc + sheen.x * sh, layer.y * sc + sheen.y * sh, layer.z * sc + sheen.z * sh);
        }
        f = v3(layer.x * (1.0f - w_cc), layer.y * (1.0f - w_cc), layer.z * (1.0f - w_cc));
        pdf += (probs[LOBE_DIFFUSE] + probs[LOBE_SHEEN]) * nl / PI_F;
        pdf += probs[LOBE_SPECULAR] * mf.pdf_reflect(v, h);
      }
      if (EXT && clearcoat > 0.0f) {
        float cnv = fmaxr(dot(ncc, v), 1.0e-4f);
        float cnl = dot(ncc, l);
        if (cnl > 0.0f) {
          float cnh = dot(ncc, h);
          float c = w_cc * d_ggx(cnh, cc_alpha) * g2_iso(cnv, cnl, cc_alpha) / (4.0f * cnv * cnl);
          f = v3(f.x + c, f.y + c, f.z + c);
          pdf += probs[LOBE_CLEARCOAT] * vndf_reflect_pdf(ncc, v, h, cc_alpha);
        }
      }
      f_out = scale(f, fmaxr(nl, 0.0f));
      pdf_out = pdf;
      return;
    }
    f_out = v3(0.0f, 0.0f, 0.0f);
    pdf_out = 0.0f;
    if (!EXT || metallic >= 1.0f || (transmission <= 0.0f && dt <= 0.0f)) return;
    float nl = -dot(n, l);
    if (nl <= 0.0f) return;
    float smax = max3(sheen);
    float layer_scale = (1.0f - w_cc) * (smax > 0.0f ? 1.0f - smax * sheen_albedo(nv, sheen_rough) : 1.0f);
    V3 f = v3(0.0f, 0.0f, 0.0f);
    float pdf = 0.0f;
    if (dt > 0.0f && transmission < 1.0f) {
      V3 fd, fm;
      fresnel_pair(nv, fd, fm);
      float w = (1.0f - metallic) * (1.0f - transmission) * dt * (1.0f - max3(fd)) * layer_scale / PI_F;
      f = v3(f.x + w * dt_color.x * nl, f.y + w * dt_color.y * nl, f.z + w * dt_color.z * nl);
      pdf += probs[LOBE_DIFFUSE_T] * nl / PI_F;
    }
    if (transmission <= 0.0f) {
      f_out = f;
      pdf_out = pdf;
      return;
    }
    float w = (1.0f - metallic) * transmission * layer_scale;
    float val, pdf_t;
    if (thin) {
      V3 lm = add(l, scale(n, 2.0f * nl));
      V3 h = normalise(add(v, lm));
      float vh = dot(v, h);
      float nh = dot(n, h);
      if (vh <= 0.0f || nh <= 0.0f) {
        f_out = f;
        pdf_out = pdf;
        return;
      }
      V3 fdv, fmv;
      fresnel_pair(vh, fdv, fmv);
      float fd = max3(fdv);
      val = mf.d(h) * mf.g2(v, lm) / (4.0f * nv * nl) * (1.0f - fd);
      pdf_t = mf.pdf_reflect(v, h);
    } else {
      float ei = eta_i, et = eta_t;
      V3 h = normalise(v3(-(ei * v.x + et * l.x), -(ei * v.y + et * l.y), -(ei * v.z + et * l.z)));
      if (dot(h, n) < 0.0f) h = neg(h);
      float vh = dot(v, h);
      float lh = dot(l, h);
      float nh = dot(n, h);
      if (vh <= 0.0f || lh >= 0.0f || nh <= 0.0f) {
        f_out = f;
        pdf_out = pdf;
        return;
      }
      float denom = ei * vh + et * lh;
      if (fabsf(denom) < 1.0e-6f) {
        f_out = f;
        pdf_out = pdf;
        return;
      }
      float d = mf.d(h);
      V3 fdv, fmv;
      fresnel_pair(vh, fdv, fmv);
      float fd = max3(fdv);
      float jac = et * et * fabsf(lh) / (denom * denom);
      val = vh * fabsf(lh) * d * mf.g2(v, l) * et * et * (1.0f - fd) / (nv * nl * denom * denom);
      pdf_t = mf.g1(v) * vh * d / nv * jac;
    }
    f = v3(f.x + base.x * w * val * nl, f.y + base.y * w * val * nl, f.z + base.z * w * val * nl);
    f_out = f;
    pdf_out = pdf + probs[LOBE_TRANSMISSION] * pdf_t;
  }

  // Bsdf::sample (pathtrace.rs:1216-1279).
  OXR_BSDF_CALL bool sample(float u0, float u1, float u2, V3& l_out, V3& f_out, float& pdf_out) const {
    float acc = 0.0f;
    int lobe = -1;
#pragma unroll
    for (int i = 0; i < (EXT ? LOBES : 2); ++i) {
      acc += probs[i];
      if (lobe < 0 && probs[i] > 0.0f && u2 < acc) lobe = i;
    }
    if (lobe < 0) {
#pragma unroll
      for (int i = 0; i < (EXT ? LOBES : 2); ++i)
        if (probs[i] > 0.0f) lobe = i;
      if (lobe < 0) return false;
    }
    V3 l;
    if (lobe == LOBE_DIFFUSE || lobe == LOBE_SHEEN || lobe == LOBE_DIFFUSE_T) {
      V3 t, b;
      basis(n, t, b);
      V3 d = cosine_hemisphere(u0, u1);
      if (lobe == LOBE_DIFFUSE_T) d.z = -d.z;
      l = to_world(t, b, n, d);
    } else if (EXT && lobe == LOBE_CLEARCOAT) {
      V3 t, b;
      basis(ncc, t, b);
      V3 vl = v3(dot(v, t), dot(v, b), dot(v, ncc));
      if (vl.z <= 0.0f) return false;
      V3 h = to_world(t, b, ncc, sample_vndf2(vl, cc_alpha, cc_alpha, u0, u1));
      l = reflect(v, h);
    } else {
      V3 h;
      if (!mf.sample(v, u0, u1, h)) return false;
      if (!EXT || lobe == LOBE_SPECULAR) {
        l = reflect(v, h);
      } else if (thin) {
        V3 r = reflect(v, h);
        float d = dot(r, n);
        l = add(r, scale(n, -2.0f * d));
      } else {
        if (!refract(v, h, eta_i / eta_t, l)) return false;
      }
    }
    if (!finite3(l) || dot(l, l) < 0.5f) return false;
    bool reflection = dot(ng, l) > 0.0f;
    bool transmits = lobe == LOBE_TRANSMISSION || lobe == LOBE_DIFFUSE_T;
    if (reflection == transmits) return false;
    V3 f;
    float pdf;
    eval(l, f, pdf);
    if (pdf <= 0.0f || !isfinite(pdf) || is_black(f) || !finite3(f)) return false;
    l_out = l;
    f_out = f;
    pdf_out = pdf;
    return true;
  }
};

using Bsdf = BsdfT<2>;

// ---- spherical triangles (pathtrace.rs:1749-1871), f64 inside ----------
struct D3 {
  double x, y, z;
};
OXR_D D3 d3sub(V3 q, V3 p) { return D3{(double)(q.x - p.x), (double)(q.y - p.y), (double)(q.z - p.z)}; }
OXR_D double d3dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
OXR_D D3 d3cross(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

OXR_D float triangle_solid_angle(V3 p, V3 a, V3 b, V3 c) {
  auto unit = [&](V3 q) -> D3 {
    D3 d = d3sub(q, p);
    double l = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
    if (l > 0.0) return D3{d.x / l, d.y / l, d.z / l};
    return D3{0.0, 0.0, 0.0};
  };
  D3 A = unit(a), B = unit(b), C = unit(c);
  D3 bc = D3{B.y * C.z - B.z * C.y, B.z * C.x - B.x * C.z, B.x * C.y - B.y * C.x};
  double num = fabs(d3dot(A, bc));
  double den = 1.0 + d3dot(A, B) + d3dot(B, C) + d3dot(C, A);
  double o = 2.0 * atan2(num, den);
  if (isfinite(o)) return (float)fmax(o, 0.0);
  return 0.0f;
}

OXR_D bool d3norm(D3 x, D3& o) {
  double l = sqrt(d3dot(x, x));
  if (!(l > 1.0e-300 && isfinite(l))) return false;
  o = D3{x.x / l, x.y / l, x.z / l};
  return true;
}

OXR_BSDF_CALL bool sample_spherical_triangle(V3 p, V3 a, V3 b, V3 c, float u0, float u1, V3& out) {
  D3 A, B, C;
  if (!d3norm(d3sub(a, p), A) || !d3norm(d3sub(b, p), B) || !d3norm(d3sub(c, p), C)) return false;
  D3 nab, nbc, nca;
  if (!d3norm(d3cross(A, B), nab) || !d3norm(d3cross(B, C), nbc) || !d3norm(d3cross(C, A), nca)) return false;
  double alpha = acos(clampd(-d3dot(nab, nca), -1.0, 1.0));
  double beta = acos(clampd(-d3dot(nbc, nab), -1.0, 1.0));
  double gamma = acos(clampd(-d3dot(nca, nbc), -1.0, 1.0));
  double area = alpha + beta + gamma - 3.141592653589793;
  if (isnan(area) || area <= 0.0) return false;
  double ap = (double)u0 * area;
  double s = sin(ap - alpha), t = cos(ap - alpha);
  double sa = sin(alpha), ca = cos(alpha);
  double cos_c = d3dot(A, B);
  double uu = t - ca;
  double vv = s + sa * cos_c;
  double den = (vv * s + uu * t) * sa;
  if (fabs(den) < 1.0e-300) return false;
  double q = clampd(((vv * t - uu * s) * ca - vv) / den, -1.0, 1.0);
  double cadot = d3dot(C, A);
  D3 ca_perp;
  if (!d3norm(D3{C.x - cadot * A.x, C.y - cadot * A.y, C.z - cadot * A.z}, ca_perp)) return false;
  double r = sqrt(fmax(1.0 - q * q, 0.0));
  D3 cp = D3{q * A.x + r * ca_perp.x, q * A.y + r * ca_perp.y, q * A.z + r * ca_perp.z};
  double z = 1.0 - (double)u1 * (1.0 - d3dot(cp, B));
  double cpb = d3dot(cp, B);
  D3 cb_perp;
  bool has_perp = d3norm(D3{cp.x - cpb * B.x, cp.y - cpb * B.y, cp.z - cpb * B.z}, cb_perp);
  double w = sqrt(fmax(1.0 - z * z, 0.0));
  D3 d = has_perp ? D3{z * B.x + w * cb_perp.x, z * B.y + w * cb_perp.y, z * B.z + w * cb_perp.z} : B;
  D3 dn;
  if (!d3norm(d, dn)) return false;
  out = v3((float)dn.x, (float)dn.y, (float)dn.z);
  return true;
}

OXR_D float power_heuristic(float a, float b) {
  float a2 = a * a, b2 = b * b;
  if (a2 + b2 > 0.0f && isfinite(a2 + b2)) return a2 / (a2 + b2);
  if (isinf(a)) return 1.0f;
  return 0.0f;
}

}  // namespace oxr
