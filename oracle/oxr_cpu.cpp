// oxr-cpu — the ORACLE of this repository: a CPU restatement of
// oxideav-render 0.0.6's PathTrace estimator (src/pathtrace.rs, the
// normative spec at :1-277, and the trace layer of src/trace.rs).
//
// This is test infrastructure, not product code. Only the tests, the
// smoke check and the bench's cpu_baseline leg load it, as the checker
// and as the timed host-CPU baseline. The product (liboxr.so) never
// links or calls it.
//
// It is written path-at-a-time, in the shape of the reference's
// `Integrator::sample`, with the reference's f32/f64 split, so a bug in
// the device wavefront cannot hide in shared code. Build:
// g++ -O2 -ffp-contract=off (no fast-math; see oracle/Makefile).
//
// Parity pins (what checks each part; see tests/test_oracle_*.py):
//   sampling (pcg/Sobol'/Owen)   Joe-Kuo table + stratification
//                                (src/pathtrace.rs:2866-2908)
//   watertight hits              brute force over all triangles, the box
//                                closed forms of src/trace.rs:1222-1241
//   offset / barycentrics        src/trace.rs:1255-1308
//   BSDF                         quadrature vs Monte Carlo, energy, pdf
//                                normalisation (src/pathtrace.rs:2954-3228)
//   spherical triangles          octant pi/2, uniformity (:3232-3296)
//   integrator                   furnace 2.0, MIS agreement, point light
//                                4/pi, clamp, glass 0.92/0.46, ambient sky
//                                0.5, determinism/additivity
//                                (:3346-3750)
//   The third-party BVH (oxideav-mesh3d 0.0.x, crates.io, not vendored)
//   is replaced by this file's own binned-SAH BVH; its watertight test
//   follows Woop, Benthin, Wald, JCGT 2(1) 2013, with the tie-break
//   (equal t -> lowest global id) defined here. Per-pixel parity against
//   the Rust reference itself is unpinned (no Rust toolchain).

#include "oxr/pathtrace.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <memory>
#include <cstring>
#include <limits>
#include <mutex>
#include <thread>
#include <vector>

namespace oc {

using f3 = std::array<float, 3>;
using d3 = std::array<double, 3>;

static const float PI_F = 3.14159265358979323846f;
static const float FRAC_PI_2_F = 1.57079632679489661923f;
static const float FRAC_PI_4_F = 0.78539816339744830962f;
static const float F32_EPS = 1.1920929e-7f;
static const float INF_F = std::numeric_limits<float>::infinity();
static const uint32_t NO_LIGHT = 0xffffffffu;

// ---------------------------------------------------------------------
// Scalar helpers with Rust's semantics.
// ---------------------------------------------------------------------

// f32::max / f32::min ignore a NaN operand.
static inline float fmaxr(float a, float b) { return std::fmax(a, b); }
static inline float fminr(float a, float b) { return std::fmin(a, b); }
// f32::clamp: NaN stays NaN.
static inline float clampr(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
static inline double clampd(double x, double lo, double hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
// powi(5) as LLVM expands it: x * ((x*x)*(x*x)).
static inline float powi5(float x) {
  float x2 = x * x;
  float x4 = x2 * x2;
  return x * x4;
}
static inline float powi2(float x) { return x * x; }
// `as usize` / `as i32`: saturating, NaN -> 0.
static inline long long sat_ll(float x) {
  if (std::isnan(x)) return 0;
  if (x >= 9.2e18f) return (long long)9.2e18;
  if (x <= -9.2e18f) return -(long long)9.2e18;
  return (long long)x;
}
static inline int32_t sat_i32(float x) {
  if (std::isnan(x)) return 0;
  if (x >= 2147483647.0f) return 2147483647;
  if (x <= -2147483648.0f) return (-2147483647 - 1);
  return (int32_t)x;
}
static inline size_t sat_usize(float x) {
  if (!(x > 0.0f)) return 0;
  if (x >= 1.8e19f) return (size_t)-1;
  return (size_t)x;
}
static inline bool finitef(float x) { return std::isfinite(x); }

// ---------------------------------------------------------------------
// Vector helpers (src/math.rs:52-90, src/pathtrace.rs:431-466).
// ---------------------------------------------------------------------
static inline f3 add(f3 a, f3 b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
static inline f3 sub(f3 a, f3 b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
static inline f3 scale(f3 a, float s) { return {a[0] * s, a[1] * s, a[2] * s}; }
static inline f3 mul(f3 a, f3 b) { return {a[0] * b[0], a[1] * b[1], a[2] * b[2]}; }
static inline f3 neg(f3 a) { return {-a[0], -a[1], -a[2]}; }
static inline float dot(f3 a, f3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline f3 cross(f3 a, f3 b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
// vec3_normalise: below f32::EPSILON the zero vector (src/math.rs:83-90).
static inline f3 normalise(f3 v) {
  float len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (len < F32_EPS) return {0.0f, 0.0f, 0.0f};
  return {v[0] / len, v[1] / len, v[2] / len};
}
static inline float max3(f3 a) { return fmaxr(fmaxr(a[0], a[1]), a[2]); }
static inline float mean3(f3 a) { return (a[0] + a[1] + a[2]) * (1.0f / 3.0f); }
static inline float luminance(f3 a) { return 0.2126f * a[0] + 0.7152f * a[1] + 0.0722f * a[2]; }
static inline bool is_black(f3 a) { return !(a[0] > 0.0f || a[1] > 0.0f || a[2] > 0.0f); }
static inline bool finite3(f3 a) { return finitef(a[0]) && finitef(a[1]) && finitef(a[2]); }

// ---------------------------------------------------------------------
// Sampling primitives (src/pathtrace.rs:297-425).
// ---------------------------------------------------------------------
uint32_t pcg_hash(uint32_t x) {
  uint32_t state = x * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
  return (word >> 22) ^ word;
}
uint32_t hash_mix(uint32_t a, uint32_t b) { return pcg_hash(a ^ pcg_hash(b)); }

struct SobolTable {
  uint32_t m[4][32];
  SobolTable() {
    for (int k = 0; k < 32; ++k) m[0][k] = 1u << (31 - k);
    // (degree s, coefficients a, initial m) — Joe & Kuo 2008,
    // src/pathtrace.rs:326-343.
    const int s_[3] = {1, 2, 3};
    const uint32_t a_[3] = {0, 1, 1};
    const uint32_t init[3][3] = {{1, 0, 0}, {1, 3, 0}, {1, 3, 1}};
    for (int d = 0; d < 3; ++d) {
      int s = s_[d];
      uint32_t a = a_[d];
      uint32_t mm[32] = {0};
      for (int k = 0; k < s; ++k) mm[k] = init[d][k];
      for (int k = s; k < 32; ++k) {
        uint32_t v = mm[k - s] ^ (mm[k - s] << s);
        for (int j = 1; j < s; ++j) {
          if (((a >> (s - 1 - j)) & 1u) == 1u) v ^= mm[k - j] << j;
        }
        mm[k] = v;
      }
      for (int k = 0; k < 32; ++k) m[d + 1][k] = mm[k] << (31 - k);
    }
  }
};
static const SobolTable& sobol_matrices() {
  static const SobolTable t;
  return t;
}

uint32_t sobol(uint32_t index, int dim) {
  const uint32_t* m = sobol_matrices().m[dim];
  uint32_t r = 0, i = index;
  int k = 0;
  while (i != 0) {
    if (i & 1u) r ^= m[k];
    i >>= 1;
    ++k;
  }
  return r;
}

static inline uint32_t reverse_bits(uint32_t x) {
  x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
  x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
  x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
  x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
  return (x >> 16) | (x << 16);
}
static inline uint32_t laine_karras(uint32_t x, uint32_t seed) {
  x += seed;
  x ^= x * 0x6c50b47cu;
  x ^= x * 0xb82f1e52u;
  x ^= x * 0xc7afe638u;
  x ^= x * 0x8d22f6e6u;
  return x;
}
static inline uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
  return reverse_bits(laine_karras(reverse_bits(x), seed));
}
static inline float to_unit(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }

std::array<float, 4> sobol_owen_4d(uint32_t index, uint32_t pattern_seed) {
  uint32_t i = nested_uniform_scramble(index, pattern_seed);
  std::array<float, 4> out;
  for (int d = 0; d < 4; ++d)
    out[d] = to_unit(nested_uniform_scramble(sobol(i, d), hash_mix(pattern_seed, (uint32_t)d)));
  return out;
}

struct Sampler {
  uint32_t pixel_seed, index, path_seed;
  Sampler(uint32_t seed, uint32_t x, uint32_t y, uint32_t idx) {
    pixel_seed = hash_mix(hash_mix(seed, y), x);
    index = idx;
    path_seed = hash_mix(pixel_seed, idx);
  }
  std::array<float, 4> pattern(uint32_t p) const {
    return sobol_owen_4d(index, hash_mix(pixel_seed, p));
  }
  float coin(uint32_t ray, uint32_t tri) const {
    return to_unit(hash_mix(hash_mix(path_seed, ray), tri));
  }
};

// ---------------------------------------------------------------------
// Frames and hemisphere sampling (src/pathtrace.rs:468-506).
// ---------------------------------------------------------------------
static void basis(f3 n, f3& t, f3& b) {
  float sign = std::copysign(1.0f, n[2]);
  float a = -1.0f / (sign + n[2]);
  float bb = n[0] * n[1] * a;
  t = {1.0f + sign * n[0] * n[0] * a, sign * bb, -sign * n[0]};
  b = {bb, sign + n[1] * n[1] * a, -n[1]};
}
static f3 to_world(f3 t, f3 b, f3 n, f3 v) {
  return {t[0] * v[0] + b[0] * v[1] + n[0] * v[2], t[1] * v[0] + b[1] * v[1] + n[1] * v[2],
          t[2] * v[0] + b[2] * v[1] + n[2] * v[2]};
}
static f3 cosine_hemisphere(float u0, float u1) {
  float a = 2.0f * u0 - 1.0f;
  float b = 2.0f * u1 - 1.0f;
  float r, phi;
  if (a == 0.0f && b == 0.0f) {
    r = 0.0f;
    phi = 0.0f;
  } else if (std::fabs(a) > std::fabs(b)) {
    r = a;
    phi = FRAC_PI_4_F * (b / a);
  } else {
    r = b;
    phi = FRAC_PI_2_F - FRAC_PI_4_F * (a / b);
  }
  float x = r * std::cos(phi), y = r * std::sin(phi);
  return {x, y, std::sqrt(fmaxr(1.0f - x * x - y * y, 0.0f))};
}

// ---------------------------------------------------------------------
// Microfacet terms (src/pathtrace.rs:512-611).
// ---------------------------------------------------------------------
static const float MIN_ALPHA = 1.0e-3f;

static inline float d_ggx(float n_dot_h, float a) {
  if (n_dot_h <= 0.0f) return 0.0f;
  float a2 = a * a;
  float d = n_dot_h * n_dot_h * (a2 - 1.0f) + 1.0f;
  return a2 / (PI_F * d * d);
}
static inline float lambda_iso(float cos, float a) {
  float c = fmaxr(std::fabs(cos), 1.0e-7f);
  float a2 = a * a;
  return (std::sqrt(a2 + (1.0f - a2) * c * c) / c - 1.0f) * 0.5f;
}
static inline float g1_iso(float cos, float a) { return 1.0f / (1.0f + lambda_iso(cos, a)); }
static inline float g2_iso(float cv, float cl, float a) {
  return 1.0f / (1.0f + lambda_iso(cv, a) + lambda_iso(cl, a));
}
static inline f3 schlick(f3 f0, float cos) {
  float k = powi5(1.0f - clampr(cos, 0.0f, 1.0f));
  return {f0[0] + (1.0f - f0[0]) * k, f0[1] + (1.0f - f0[1]) * k, f0[2] + (1.0f - f0[2]) * k};
}
static f3 sample_vndf(f3 v, float a, float u0, float u1) {
  f3 vh = normalise({a * v[0], a * v[1], v[2]});
  float lensq = vh[0] * vh[0] + vh[1] * vh[1];
  f3 t1;
  if (lensq > 0.0f) {
    float il = 1.0f / std::sqrt(lensq);
    t1 = {-vh[1] * il, vh[0] * il, 0.0f};
  } else {
    t1 = {1.0f, 0.0f, 0.0f};
  }
  f3 t2 = cross(vh, t1);
  float r = std::sqrt(u0);
  float phi = 2.0f * PI_F * u1;
  float p1 = r * std::cos(phi);
  float p2 = r * std::sin(phi);
  float s = 0.5f * (1.0f + vh[2]);
  p2 = (1.0f - s) * std::sqrt(fmaxr(1.0f - p1 * p1, 0.0f)) + s * p2;
  float pz = std::sqrt(fmaxr(1.0f - p1 * p1 - p2 * p2, 0.0f));
  f3 nh = {p1 * t1[0] + p2 * t2[0] + pz * vh[0], p1 * t1[1] + p2 * t2[1] + pz * vh[1],
           p1 * t1[2] + p2 * t2[2] + pz * vh[2]};
  return normalise({a * nh[0], a * nh[1], fmaxr(nh[2], 1.0e-6f)});
}
static f3 sample_vndf_aniso(f3 v, float at, float ab, float u0, float u1) {
  f3 vh = normalise({at * v[0], ab * v[1], v[2]});
  float lensq = vh[0] * vh[0] + vh[1] * vh[1];
  f3 t1;
  if (lensq > 0.0f) {
    float il = 1.0f / std::sqrt(lensq);
    t1 = {-vh[1] * il, vh[0] * il, 0.0f};
  } else {
    t1 = {1.0f, 0.0f, 0.0f};
  }
  f3 t2 = cross(vh, t1);
  float r = std::sqrt(u0);
  float phi = 2.0f * PI_F * u1;
  float p1 = r * std::cos(phi);
  float p2 = r * std::sin(phi);
  float s = 0.5f * (1.0f + vh[2]);
  p2 = (1.0f - s) * std::sqrt(fmaxr(1.0f - p1 * p1, 0.0f)) + s * p2;
  float pz = std::sqrt(fmaxr(1.0f - p1 * p1 - p2 * p2, 0.0f));
  f3 nh = {p1 * t1[0] + p2 * t2[0] + pz * vh[0], p1 * t1[1] + p2 * t2[1] + pz * vh[1],
           p1 * t1[2] + p2 * t2[2] + pz * vh[2]};
  return normalise({at * nh[0], ab * nh[1], fmaxr(nh[2], 1.0e-6f)});
}
static inline float vndf_reflect_pdf(f3 n, f3 v, f3 h, float a) {
  float nv = dot(n, v);
  float nh = dot(n, h);
  if (nv <= 0.0f || nh <= 0.0f || dot(v, h) <= 0.0f) return 0.0f;
  return g1_iso(nv, a) * d_ggx(nh, a) / (4.0f * nv);
}
static inline f3 reflect(f3 v, f3 h) {
  float d = 2.0f * dot(v, h);
  return {d * h[0] - v[0], d * h[1] - v[1], d * h[2] - v[2]};
}
static bool refract(f3 v, f3 h, float eta, f3& out) {
  float ci = dot(v, h);
  float s2 = eta * eta * fmaxr(1.0f - ci * ci, 0.0f);
  if (s2 >= 1.0f) return false;
  float ct = std::sqrt(1.0f - s2);
  float k = eta * ci - ct;
  out = normalise({-eta * v[0] + k * h[0], -eta * v[1] + k * h[1], -eta * v[2] + k * h[2]});
  return true;
}

// ---------------------------------------------------------------------
// Sheen (src/pathtrace.rs:617-692).
// ---------------------------------------------------------------------
static const int SHEEN_LUT_COS = 32;
static const int SHEEN_LUT_ROUGH = 16;

static inline float d_charlie(float n_dot_h, float a) {
  float inv = 1.0f / fmaxr(a, 1.0e-3f);
  float sin2 = fmaxr(1.0f - n_dot_h * n_dot_h, 0.0f);
  return (2.0f + inv) * std::pow(sin2, inv * 0.5f) / (2.0f * PI_F);
}
static inline float v_neubelt(float nl, float nv) {
  return 1.0f / fmaxr(4.0f * (nl + nv - nl * nv), 1.0e-6f);
}
struct SheenLut {
  float v[SHEEN_LUT_COS * SHEEN_LUT_ROUGH];
  SheenLut() {
    const int nt = 48, np = 24;
    for (int ri = 0; ri < SHEEN_LUT_ROUGH; ++ri) {
      float r = (float)ri / (float)(SHEEN_LUT_ROUGH - 1);
      float a = fmaxr(r * r, 1.0e-3f);
      for (int ci = 0; ci < SHEEN_LUT_COS; ++ci) {
        float cv = fmaxr(((float)ci + 0.5f) / (float)SHEEN_LUT_COS, 1.0e-3f);
        f3 vv = {std::sqrt(1.0f - cv * cv), 0.0f, cv};
        float sum = 0.0f;
        for (int i = 0; i < nt; ++i) {
          float th = ((float)i + 0.5f) / (float)nt * FRAC_PI_2_F;
          float st = std::sin(th), ct = std::cos(th);
          for (int j = 0; j < np; ++j) {
            float ph = ((float)j + 0.5f) / (float)np * PI_F;
            f3 l = {st * std::cos(ph), st * std::sin(ph), ct};
            f3 h = normalise(add(vv, l));
            float f = d_charlie(h[2], a) * v_neubelt(ct, cv);
            sum += f * ct * st;
          }
        }
        float dw = (FRAC_PI_2_F / (float)nt) * (PI_F / (float)np) * 2.0f;
        v[ri * SHEEN_LUT_COS + ci] = clampr(sum * dw, 0.0f, 1.0f);
      }
    }
  }
};
static const SheenLut& sheen_lut() {
  static const SheenLut l;
  return l;
}
static float sheen_albedo(float cos, float rough) {
  const float* lut = sheen_lut().v;
  float x = clampr(clampr(cos, 0.0f, 1.0f) * (float)SHEEN_LUT_COS - 0.5f, 0.0f,
                   (float)(SHEEN_LUT_COS - 1));
  float y = clampr(clampr(rough, 0.0f, 1.0f) * (float)(SHEEN_LUT_ROUGH - 1), 0.0f,
                   (float)(SHEEN_LUT_ROUGH - 1));
  size_t x0 = sat_usize(std::floor(x)), y0 = sat_usize(std::floor(y));
  size_t x1 = std::min<size_t>(x0 + 1, SHEEN_LUT_COS - 1);
  size_t y1 = std::min<size_t>(y0 + 1, SHEEN_LUT_ROUGH - 1);
  float fx = x - (float)x0, fy = y - (float)y0;
  auto at = [&](size_t xx, size_t yy) { return lut[yy * SHEEN_LUT_COS + xx]; };
  float a = at(x0, y0) + (at(x1, y0) - at(x0, y0)) * fx;
  float b = at(x0, y1) + (at(x1, y1) - at(x0, y1)) * fx;
  return a + (b - a) * fy;
}

// ---------------------------------------------------------------------
// Thin-film iridescence (src/pathtrace.rs:1326-1425).
// ---------------------------------------------------------------------
static f3 eval_sensitivity(float opd, f3 shift) {
  float phase = 2.0f * PI_F * opd * 1.0e-9f;
  const float val[3] = {5.4856e-13f, 4.4201e-13f, 5.2481e-13f};
  const float pos[3] = {1.6810e6f, 1.7953e6f, 2.2084e6f};
  const float var[3] = {4.3278e9f, 9.3046e9f, 6.6121e9f};
  float xyz[3];
  for (int k = 0; k < 3; ++k)
    xyz[k] = val[k] * std::sqrt(2.0f * PI_F * var[k]) * std::cos(pos[k] * phase + shift[k]) *
             std::exp(-phase * phase * var[k]);
  xyz[0] += 9.7470e-14f * std::sqrt(2.0f * PI_F * 4.5282e9f) * std::cos(2.2399e6f * phase + shift[0]) *
            std::exp(-4.5282e9f * phase * phase);
  for (int k = 0; k < 3; ++k) xyz[k] /= 1.0685e-7f;
  return {3.2404542f * xyz[0] - 1.5371385f * xyz[1] - 0.4985314f * xyz[2],
          -0.969266f * xyz[0] + 1.8760108f * xyz[1] + 0.041556f * xyz[2],
          0.0556434f * xyz[0] - 0.2040259f * xyz[1] + 1.0572252f * xyz[2]};
}

f3 iridescent_fresnel(float outside_ior, float film_ior, f3 base_f0, float thickness_nm, float cos1) {
  cos1 = clampr(cos1, 0.0f, 1.0f);
  float r0 = powi2((film_ior - outside_ior) / (film_ior + outside_ior));
  float r12 = r0 + (1.0f - r0) * powi5(1.0f - cos1);
  float t121 = 1.0f - r12;
  float sin2_2 = powi2(outside_ior / film_ior) * (1.0f - cos1 * cos1);
  float cos2_sq = 1.0f - sin2_2;
  if (cos2_sq < 0.0f) return {1.0f, 1.0f, 1.0f};
  float cos2 = std::sqrt(cos2_sq);
  float r23[3], phi23[3] = {0.0f, 0.0f, 0.0f};
  for (int k = 0; k < 3; ++k) {
    float f0 = clampr(base_f0[k] + 1.0e-4f, 0.0f, 0.9999f);
    float sq = std::sqrt(f0);
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
  for (int k = 0; k < 3; ++k) {
    r123[k] = clampr(r12 * r23[k], 1.0e-5f, 0.9999f);
    rs[k] = t121 * t121 * r23[k] / (1.0f - r123[k]);
    out[k] = r12 + rs[k];
  }
  float cm[3] = {rs[0] - t121, rs[1] - t121, rs[2] - t121};
  for (int m = 1; m <= 2; ++m) {
    for (int k = 0; k < 3; ++k) cm[k] *= std::sqrt(r123[k]);
    f3 sm = eval_sensitivity((float)m * opd, {(float)m * phi[0], (float)m * phi[1], (float)m * phi[2]});
    for (int k = 0; k < 3; ++k) out[k] += cm[k] * 2.0f * sm[k];
  }
  return {clampr(out[0], 0.0f, 1.0f), clampr(out[1], 0.0f, 1.0f), clampr(out[2], 0.0f, 1.0f)};
}

static f3 dispersion_iors(float ior, float dispersion) {
  float hs = (ior - 1.0f) * 0.025f * fmaxr(dispersion, 0.0f);
  return {fmaxr(ior - hs, 1.0f), fmaxr(ior, 1.0f), fmaxr(ior + hs, 1.0f)};
}

// ---------------------------------------------------------------------
// MaterialSample (src/trace.rs:201-305) and the layered BSDF
// (src/pathtrace.rs:706-1280).
// ---------------------------------------------------------------------
struct MaterialSample {
  float base_color[4] = {1, 1, 1, 1};
  float metallic = 1.0f, roughness = 1.0f;
  f3 normal = {0, 0, 1};
  f3 emissive = {0, 0, 0};
  bool unlit = false;
  bool double_sided = false;
  float ior = 1.5f;
  float specular = 1.0f;
  f3 specular_color = {1, 1, 1};
  float transmission = 0.0f;
  float thickness = 0.0f;
  f3 attenuation_color = {1, 1, 1};
  float attenuation_distance = INF_F;
  float clearcoat = 0.0f, clearcoat_roughness = 0.0f;
  f3 clearcoat_normal = {0, 0, 1};
  f3 sheen_color = {0, 0, 0};
  float sheen_roughness = 0.0f;
  float anisotropy = 0.0f;
  f3 anisotropy_direction = {0, 0, 0};
  float iridescence = 0.0f, iridescence_ior = 1.3f, iridescence_thickness = 400.0f;
  float diffuse_transmission = 0.0f;
  f3 diffuse_transmission_color = {1, 1, 1};
  float dispersion = 0.0f;
};

enum { LOBE_DIFFUSE = 0, LOBE_SPECULAR, LOBE_TRANSMISSION, LOBE_CLEARCOAT, LOBE_SHEEN, LOBE_DIFFUSE_T, LOBES };

struct Microfacet {
  f3 t, b, n;
  float at, ab;
  bool aniso;
  static Microfacet isotropic(f3 n, float a) {
    Microfacet m;
    basis(n, m.t, m.b);
    m.n = n;
    m.at = a;
    m.ab = a;
    m.aniso = false;
    return m;
  }
  f3 local(f3 w) const { return {dot(w, t), dot(w, b), dot(w, n)}; }
  float d(f3 h) const {
    if (!aniso) return d_ggx(dot(n, h), at);
    f3 hl = local(h);
    if (hl[2] <= 0.0f) return 0.0f;
    float x = hl[0] / at;
    float y = hl[1] / ab;
    float e = x * x + y * y + hl[2] * hl[2];
    return 1.0f / (PI_F * at * ab * e * e);
  }
  float lambda(f3 w) const {
    if (!aniso) return lambda_iso(dot(n, w), at);
    f3 wl = local(w);
    float cz = fmaxr(std::fabs(wl[2]), 1.0e-7f);
    float a2 = powi2(at * wl[0]) + powi2(ab * wl[1]);
    return (std::sqrt(1.0f + a2 / (cz * cz)) - 1.0f) * 0.5f;
  }
  float g1(f3 v) const { return 1.0f / (1.0f + lambda(v)); }
  float g2(f3 v, f3 l) const { return 1.0f / (1.0f + lambda(v) + lambda(l)); }
  float pdf_reflect(f3 v, f3 h) const {
    if (!aniso) return vndf_reflect_pdf(n, v, h, at);
    float nv = dot(n, v);
    float nh = dot(n, h);
    if (nv <= 0.0f || nh <= 0.0f || dot(v, h) <= 0.0f) return 0.0f;
    return g1(v) * d(h) / (4.0f * nv);
  }
  bool sample(f3 v, float u0, float u1, f3& out) const {
    f3 vl = local(v);
    if (vl[2] <= 0.0f) return false;
    f3 h = aniso ? sample_vndf_aniso(vl, at, ab, u0, u1) : sample_vndf(vl, at, u0, u1);
    out = to_world(t, b, n, h);
    return true;
  }
};

struct Bsdf {
  f3 v, n, ng, ncc;
  f3 base;
  float metallic;
  Microfacet mf;
  float specular;
  f3 f0d;
  float transmission;
  bool thin;
  float eta_i, eta_t;
  float clearcoat, cc_alpha;
  f3 sheen;
  float sheen_rough, sheen_alpha;
  float irid, irid_ior, irid_thick;
  float dt;
  f3 dt_color;
  float w_cc;
  float probs[LOBES];

  // Bsdf::new (src/pathtrace.rs:853-986).
  Bsdf(const MaterialSample& m, f3 v_, f3 ng_in, bool front) {
    auto flip = [&](f3 a) { return front ? a : neg(a); };
    v = v_;
    ng = flip(ng_in);
    n = flip(m.normal);
    if (dot(n, v) <= 1.0e-4f) n = ng;
    ncc = flip(m.clearcoat_normal);
    if (dot(ncc, v) <= 1.0e-4f) ncc = n;
    thin = m.thickness <= 0.0f;
    float ior = (finitef(m.ior) && m.ior >= 1.0f) ? m.ior : 1.5f;
    if (thin || front) {
      eta_i = 1.0f;
      eta_t = ior;
    } else {
      eta_i = ior;
      eta_t = 1.0f;
    }
    bool inside = !thin && !front;
    float r0 = powi2((ior - 1.0f) / (ior + 1.0f));
    base = {m.base_color[0], m.base_color[1], m.base_color[2]};
    float alpha = fmaxr(m.roughness * m.roughness, MIN_ALPHA);
    float aniso = clampr(m.anisotropy, 0.0f, 1.0f);
    mf = Microfacet::isotropic(n, alpha);
    if (aniso > 0.0f) {
      f3 d0 = m.anisotropy_direction;
      float dn = dot(d0, n);
      f3 t = normalise({d0[0] - n[0] * dn, d0[1] - n[1] * dn, d0[2] - n[2] * dn});
      if (dot(t, t) > 0.5f) {
        float at = alpha + (1.0f - alpha) * aniso * aniso;
        if (at > alpha) {
          mf.t = t;
          mf.b = cross(n, t);
          mf.n = n;
          mf.at = at;
          mf.ab = alpha;
          mf.aniso = true;
        }
      }
    }
    metallic = m.metallic;
    specular = clampr(m.specular, 0.0f, 1.0f);
    f0d = {fminr(r0 * m.specular_color[0], 1.0f), fminr(r0 * m.specular_color[1], 1.0f),
           fminr(r0 * m.specular_color[2], 1.0f)};
    transmission = m.transmission;
    clearcoat = inside ? 0.0f : m.clearcoat;
    cc_alpha = fmaxr(m.clearcoat_roughness * m.clearcoat_roughness, MIN_ALPHA);
    sheen = inside ? f3{0, 0, 0} : m.sheen_color;
    sheen_rough = m.sheen_roughness;
    sheen_alpha = fmaxr(m.sheen_roughness * m.sheen_roughness, 1.0e-3f);
    irid = inside ? 0.0f : clampr(m.iridescence, 0.0f, 1.0f);
    irid_ior = finitef(m.iridescence_ior) ? fmaxr(m.iridescence_ior, 1.0f) : 1.3f;
    irid_thick = finitef(m.iridescence_thickness) ? fmaxr(m.iridescence_thickness, 0.0f) : 400.0f;
    dt = clampr(m.diffuse_transmission, 0.0f, 1.0f);
    dt_color = {clampr(m.diffuse_transmission_color[0], 0.0f, 1.0f),
                clampr(m.diffuse_transmission_color[1], 0.0f, 1.0f),
                clampr(m.diffuse_transmission_color[2], 0.0f, 1.0f)};
    w_cc = 0.0f;
    float cv = fmaxr(dot(n, v), 1.0e-4f);
    w_cc = clearcoat > 0.0f ? clearcoat * schlick({0.04f, 0.04f, 0.04f}, dot(ncc, v))[0] : 0.0f;
    float sheen_max = max3(sheen);
    float e_v = sheen_max > 0.0f ? sheen_albedo(cv, sheen_rough) : 0.0f;
    float s_scale = (1.0f - w_cc) * (1.0f - sheen_max * e_v);
    f3 fd_v, fm_v;
    fresnel_pair(cv, fd_v, fm_v);
    float fd = max3(fd_v);
    float fm = mean3(fm_v);
    float m_ = metallic;
    float base_mean = fmaxr(mean3(base), 0.0f);
    float t = transmission;
    float p[LOBES] = {
        s_scale * (1.0f - m_) * (1.0f - t) * (1.0f - dt) * (1.0f - fd) * base_mean,
        s_scale * ((1.0f - m_) * fd + m_ * fm),
        s_scale * (1.0f - m_) * t * (1.0f - fd) * base_mean,
        w_cc,
        (1.0f - w_cc) * sheen_max * e_v,
        s_scale * (1.0f - m_) * (1.0f - t) * dt * (1.0f - fd) * fmaxr(mean3(dt_color), 0.0f),
    };
    float sum = 0.0f;
    for (int i = 0; i < LOBES; ++i) sum += p[i];
    if (sum > 0.0f && finitef(sum)) {
      for (int i = 0; i < LOBES; ++i) probs[i] = p[i] / sum;
    } else {
      for (int i = 0; i < LOBES; ++i) probs[i] = 0.0f;
    }
  }

  bool black() const {
    for (int i = 0; i < LOBES; ++i)
      if (probs[i] > 0.0f) return false;
    return true;
  }

  f3 fresnel_d(float c) const {
    if (eta_i > eta_t) {
      float eta = eta_i / eta_t;
      float s2 = eta * eta * fmaxr(1.0f - c * c, 0.0f);
      if (s2 >= 1.0f) return {1.0f, 1.0f, 1.0f};
      return schlick(f0d, std::sqrt(1.0f - s2));
    }
    return schlick(f0d, c);
  }

  void fresnel_pair(float c, f3& fd_out, f3& fm_out) const {
    f3 fd = scale(fresnel_d(c), specular);
    f3 fm = schlick(base, c);
    if (irid <= 0.0f) {
      fd_out = fd;
      fm_out = fm;
      return;
    }
    float i = irid;
    auto mix3 = [&](f3 a, f3 b) {
      return f3{a[0] + (b[0] - a[0]) * i, a[1] + (b[1] - a[1]) * i, a[2] + (b[2] - a[2]) * i};
    };
    f3 fd_i = metallic < 1.0f ? iridescent_fresnel(1.0f, irid_ior, scale(f0d, specular), irid_thick, c) : fd;
    f3 fm_i = metallic > 0.0f ? iridescent_fresnel(1.0f, irid_ior, base, irid_thick, c) : fm;
    fd_out = mix3(fd, fd_i);
    fm_out = mix3(fm, fm_i);
  }

  // Bsdf::eval (src/pathtrace.rs:1075-1211): f·|n·l| and mixture pdf.
  void eval(f3 l, f3& f_out, float& pdf_out) const {
    float nv = fmaxr(dot(n, v), 1.0e-4f);
    const float* p = probs;
    if (dot(ng, l) > 0.0f) {
      float nl = dot(n, l);
      f3 h = normalise(add(v, l));
      float vh = fmaxr(dot(v, h), 0.0f);
      f3 f = {0, 0, 0};
      float pdf = 0.0f;
      if (nl > 0.0f) {
        f3 fd, fm;
        fresnel_pair(vh, fd, fm);
        float spec_c = mf.d(h) * mf.g2(v, l) / (4.0f * nv * nl);
        float m = metallic;
        float diff_k = (1.0f - m) * (1.0f - transmission) * (1.0f - dt) * (1.0f - max3(fd)) / PI_F;
        f3 layer;
        for (int k = 0; k < 3; ++k) layer[k] = spec_c * ((1.0f - m) * fd[k] + m * fm[k]) + diff_k * base[k];
        float smax = max3(sheen);
        if (smax > 0.0f) {
          float sc = fminr(1.0f - smax * sheen_albedo(nv, sheen_rough), 1.0f - smax * sheen_albedo(nl, sheen_rough));
          float nh = dot(n, h);
          float sh = d_charlie(nh, sheen_alpha) * v_neubelt(nl, nv);
          for (int k = 0; k < 3; ++k) layer[k] = layer[k] * sc + sheen[k] * sh;
        }
        for (int k = 0; k < 3; ++k) f[k] = layer[k] * (1.0f - w_cc);
        pdf += (p[LOBE_DIFFUSE] + p[LOBE_SHEEN]) * nl / PI_F;
        pdf += p[LOBE_SPECULAR] * mf.pdf_reflect(v, h);
      }
      if (clearcoat > 0.0f) {
        float cnv = fmaxr(dot(ncc, v), 1.0e-4f);
        float cnl = dot(ncc, l);
        if (cnl > 0.0f) {
          float cnh = dot(ncc, h);
          float c = w_cc * d_ggx(cnh, cc_alpha) * g2_iso(cnv, cnl, cc_alpha) / (4.0f * cnv * cnl);
          for (int k = 0; k < 3; ++k) f[k] += c;
          pdf += p[LOBE_CLEARCOAT] * vndf_reflect_pdf(ncc, v, h, cc_alpha);
        }
      }
      float cos = fmaxr(nl, 0.0f);
      f_out = scale(f, cos);
      pdf_out = pdf;
      return;
    }
    // Transmission.
    f_out = {0, 0, 0};
    pdf_out = 0.0f;
    if (metallic >= 1.0f || (transmission <= 0.0f && dt <= 0.0f)) return;
    float nl = -dot(n, l);
    if (nl <= 0.0f) return;
    float smax = max3(sheen);
    float layer_scale = (1.0f - w_cc) * (smax > 0.0f ? 1.0f - smax * sheen_albedo(nv, sheen_rough) : 1.0f);
    f3 f = {0, 0, 0};
    float pdf = 0.0f;
    if (dt > 0.0f && transmission < 1.0f) {
      f3 fd, fm;
      fresnel_pair(nv, fd, fm);
      float w = (1.0f - metallic) * (1.0f - transmission) * dt * (1.0f - max3(fd)) * layer_scale / PI_F;
      for (int k = 0; k < 3; ++k) f[k] += w * dt_color[k] * nl;
      pdf += p[LOBE_DIFFUSE_T] * nl / PI_F;
    }
    if (transmission <= 0.0f) {
      f_out = f;
      pdf_out = pdf;
      return;
    }
    float w = (1.0f - metallic) * transmission * layer_scale;
    float val, pdf_t;
    if (thin) {
      f3 lm = add(l, scale(n, 2.0f * nl));
      f3 h = normalise(add(v, lm));
      float vh = dot(v, h);
      float nh = dot(n, h);
      if (vh <= 0.0f || nh <= 0.0f) {
        f_out = f;
        pdf_out = pdf;
        return;
      }
      f3 fdv, fmv;
      fresnel_pair(vh, fdv, fmv);
      float fd = max3(fdv);
      val = mf.d(h) * mf.g2(v, lm) / (4.0f * nv * nl) * (1.0f - fd);
      pdf_t = mf.pdf_reflect(v, h);
    } else {
      float ei = eta_i, et = eta_t;
      f3 h = normalise({-(ei * v[0] + et * l[0]), -(ei * v[1] + et * l[1]), -(ei * v[2] + et * l[2])});
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
      if (std::fabs(denom) < 1.0e-6f) {
        f_out = f;
        pdf_out = pdf;
        return;
      }
      float d = mf.d(h);
      f3 fdv, fmv;
      fresnel_pair(vh, fdv, fmv);
      float fd = max3(fdv);
      float jac = et * et * std::fabs(lh) / (denom * denom);
      val = vh * std::fabs(lh) * d * mf.g2(v, l) * et * et * (1.0f - fd) / (nv * nl * denom * denom);
      pdf_t = mf.g1(v) * vh * d / nv * jac;
    }
    for (int k = 0; k < 3; ++k) f[k] += base[k] * w * val * nl;
    f_out = f;
    pdf_out = pdf + p[LOBE_TRANSMISSION] * pdf_t;
  }

  // Bsdf::sample (src/pathtrace.rs:1216-1279).
  bool sample(const float u[3], f3& l_out, f3& f_out, float& pdf_out) const {
    float acc = 0.0f;
    int lobe = -1;
    for (int i = 0; i < LOBES; ++i) {
      acc += probs[i];
      if (probs[i] > 0.0f && u[2] < acc) {
        lobe = i;
        break;
      }
    }
    if (lobe < 0) {
      for (int i = LOBES - 1; i >= 0; --i)
        if (probs[i] > 0.0f) {
          lobe = i;
          break;
        }
      if (lobe < 0) return false;
    }
    f3 l;
    switch (lobe) {
      case LOBE_DIFFUSE:
      case LOBE_SHEEN: {
        f3 t, b;
        basis(n, t, b);
        l = to_world(t, b, n, cosine_hemisphere(u[0], u[1]));
        break;
      }
      case LOBE_DIFFUSE_T: {
        f3 t, b;
        basis(n, t, b);
        f3 d = cosine_hemisphere(u[0], u[1]);
        l = to_world(t, b, n, {d[0], d[1], -d[2]});
        break;
      }
      case LOBE_CLEARCOAT: {
        f3 t, b;
        basis(ncc, t, b);
        f3 vl = {dot(v, t), dot(v, b), dot(v, ncc)};
        if (vl[2] <= 0.0f) return false;
        f3 h = to_world(t, b, ncc, sample_vndf(vl, cc_alpha, u[0], u[1]));
        l = reflect(v, h);
        break;
      }
      case LOBE_SPECULAR:
      case LOBE_TRANSMISSION: {
        f3 h;
        if (!mf.sample(v, u[0], u[1], h)) return false;
        if (lobe == LOBE_SPECULAR) {
          l = reflect(v, h);
        } else if (thin) {
          f3 r = reflect(v, h);
          float d = dot(r, n);
          l = add(r, scale(n, -2.0f * d));
        } else {
          if (!refract(v, h, eta_i / eta_t, l)) return false;
        }
        break;
      }
      default:
        return false;
    }
    if (!finite3(l) || dot(l, l) < 0.5f) return false;
    bool reflection = dot(ng, l) > 0.0f;
    bool transmits = lobe == LOBE_TRANSMISSION || lobe == LOBE_DIFFUSE_T;
    if (reflection == transmits) return false;
    f3 f;
    float pdf;
    eval(l, f, pdf);
    if (pdf <= 0.0f || !finitef(pdf) || is_black(f) || !finite3(f)) return false;
    l_out = l;
    f_out = f;
    pdf_out = pdf;
    return true;
  }
};

// ---------------------------------------------------------------------
// Spherical triangles (src/pathtrace.rs:1749-1871), f64 inside.
// ---------------------------------------------------------------------
static const float MIN_SPHERICAL_SOLID_ANGLE = 1.0e-4f;

float triangle_solid_angle(f3 p, f3 a, f3 b, f3 c) {
  auto u = [&](f3 q) -> d3 {
    d3 d = {(double)(q[0] - p[0]), (double)(q[1] - p[1]), (double)(q[2] - p[2])};
    double l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (l > 0.0) return {d[0] / l, d[1] / l, d[2] / l};
    return {0.0, 0.0, 0.0};
  };
  d3 A = u(a), B = u(b), C = u(c);
  auto dt = [](d3 x, d3 y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
  d3 bc = {B[1] * C[2] - B[2] * C[1], B[2] * C[0] - B[0] * C[2], B[0] * C[1] - B[1] * C[0]};
  double num = std::fabs(dt(A, bc));
  double den = 1.0 + dt(A, B) + dt(B, C) + dt(C, A);
  double o = 2.0 * std::atan2(num, den);
  if (std::isfinite(o)) return (float)std::fmax(o, 0.0);
  return 0.0f;
}

bool sample_spherical_triangle(f3 p, f3 a, f3 b, f3 c, float u0, float u1, f3& out) {
  auto subp = [&](f3 x) -> d3 {
    return {(double)(x[0] - p[0]), (double)(x[1] - p[1]), (double)(x[2] - p[2])};
  };
  auto dt = [](d3 x, d3 y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
  auto cr = [](d3 x, d3 y) -> d3 {
    return {x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]};
  };
  auto norm = [&](d3 x, d3& o) -> bool {
    double l = std::sqrt(dt(x, x));
    if (!(l > 1.0e-300 && std::isfinite(l))) return false;
    o = {x[0] / l, x[1] / l, x[2] / l};
    return true;
  };
  d3 A, B, C;
  if (!norm(subp(a), A) || !norm(subp(b), B) || !norm(subp(c), C)) return false;
  d3 nab, nbc, nca;
  if (!norm(cr(A, B), nab) || !norm(cr(B, C), nbc) || !norm(cr(C, A), nca)) return false;
  auto ang = [&](d3 x, d3 y) { return std::acos(clampd(-dt(x, y), -1.0, 1.0)); };
  double alpha = ang(nab, nca);
  double beta = ang(nbc, nab);
  double gamma = ang(nca, nbc);
  double area = alpha + beta + gamma - M_PI;
  if (std::isnan(area) || area <= 0.0) return false;
  double ap = (double)u0 * area;
  double s = std::sin(ap - alpha), t = std::cos(ap - alpha);
  double sa = std::sin(alpha), ca = std::cos(alpha);
  double cos_c = dt(A, B);
  double uu = t - ca;
  double vv = s + sa * cos_c;
  double den = (vv * s + uu * t) * sa;
  if (std::fabs(den) < 1.0e-300) return false;
  double q = clampd(((vv * t - uu * s) * ca - vv) / den, -1.0, 1.0);
  d3 ca_perp;
  double cadot = dt(C, A);
  if (!norm({C[0] - cadot * A[0], C[1] - cadot * A[1], C[2] - cadot * A[2]}, ca_perp)) return false;
  double r = std::sqrt(std::fmax(1.0 - q * q, 0.0));
  d3 cp = {q * A[0] + r * ca_perp[0], q * A[1] + r * ca_perp[1], q * A[2] + r * ca_perp[2]};
  double z = 1.0 - (double)u1 * (1.0 - dt(cp, B));
  double cpb = dt(cp, B);
  d3 cb_perp;
  bool has_perp = norm({cp[0] - cpb * B[0], cp[1] - cpb * B[1], cp[2] - cpb * B[2]}, cb_perp);
  double w = std::sqrt(std::fmax(1.0 - z * z, 0.0));
  d3 d = has_perp ? d3{z * B[0] + w * cb_perp[0], z * B[1] + w * cb_perp[1], z * B[2] + w * cb_perp[2]} : B;
  d3 dn;
  if (!norm(d, dn)) return false;
  out = {(float)dn[0], (float)dn[1], (float)dn[2]};
  return true;
}

// ---------------------------------------------------------------------
// Trace-layer helpers (src/trace.rs:1120-1204).
// ---------------------------------------------------------------------
bool barycentric_of(f3 q, f3 p0, f3 p1, f3 p2, f3& out) {
  f3 e1 = sub(p1, p0), e2 = sub(p2, p0), r = sub(q, p0);
  float d11 = dot(e1, e1), d12 = dot(e1, e2), d22 = dot(e2, e2);
  float r1 = dot(r, e1), r2 = dot(r, e2);
  float det = d11 * d22 - d12 * d12;
  if (!finitef(det) || std::fabs(det) <= std::numeric_limits<float>::min()) return false;
  float u = (d22 * r1 - d12 * r2) / det;
  float v = (d11 * r2 - d12 * r1) / det;
  out = {1.0f - u - v, u, v};
  return true;
}

static inline float bits_f(uint32_t b) {
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}
static inline uint32_t f_bits(float f) {
  uint32_t b;
  std::memcpy(&b, &f, 4);
  return b;
}

f3 offset_ray_origin(f3 p, f3 n) {
  const float ORIGIN = 1.0f / 32.0f;
  const float FLOAT_SCALE = 1.0f / 65536.0f;
  const float INT_SCALE = 256.0f;
  f3 out;
  for (int k = 0; k < 3; ++k) {
    int32_t of_i = sat_i32(INT_SCALE * n[k]);
    int32_t bits = (int32_t)f_bits(p[k]);
    int32_t moved = p[k] < 0.0f ? (int32_t)((uint32_t)bits - (uint32_t)of_i)
                                : (int32_t)((uint32_t)bits + (uint32_t)of_i);
    float p_i = bits_f((uint32_t)moved);
    if (std::fabs(p[k]) < ORIGIN)
      out[k] = p[k] + FLOAT_SCALE * n[k];
    else if (finitef(p_i))
      out[k] = p_i;
    else
      out[k] = p[k];
  }
  return out;
}

static inline f3 interp3(const float* v9, f3 b) {
  return {v9[0] * b[0] + v9[3] * b[1] + v9[6] * b[2], v9[1] * b[0] + v9[4] * b[1] + v9[7] * b[2],
          v9[2] * b[0] + v9[5] * b[1] + v9[8] * b[2]};
}

// ---------------------------------------------------------------------
// Scene: the prepared soup + materials + lights, and its BVH.
// ---------------------------------------------------------------------
struct PMaterial {  // PreparedMaterial + material_oriented constants
  float base_color[4];
  float metallic, roughness;
  f3 emissive;
  bool double_sided, unlit;
  uint32_t alpha_mode;
  float alpha_cutoff;
  float ior;
  float specular;
  f3 specular_color;
  float transmission, thickness;
  f3 attenuation_color;
  float attenuation_distance;
  float clearcoat, clearcoat_roughness;
  f3 sheen_color;
  float sheen_roughness;
  float iridescence, iridescence_ior, iridescence_thickness;
  float diffuse_transmission;
  f3 diffuse_transmission_color;
  float dispersion;
  float anisotropy, anisotropy_rotation;  // KHR_materials_anisotropy
};

// PreparedMaterial::from_material (src/prepare.rs:623-668) followed by
// the untextured reads of TraceScene::material_oriented
// (src/trace.rs:724-922).
static PMaterial prepare_material(const oxr_material& m) {
  auto fin = [](float v, float d) { return std::isfinite(v) ? v : d; };
  PMaterial p;
  p.base_color[0] = fin(m.base_color[0], 1.0f);
  p.base_color[1] = fin(m.base_color[1], 1.0f);
  p.base_color[2] = fin(m.base_color[2], 1.0f);
  p.base_color[3] = clampr(fin(m.base_color[3], 1.0f), 0.0f, 1.0f);
  p.metallic = clampr(fin(m.metallic, 1.0f), 0.0f, 1.0f);
  p.roughness = clampr(fin(m.roughness, 1.0f), 0.0f, 1.0f);
  for (int k = 0; k < 3; ++k) p.emissive[k] = fmaxr(fin(m.emissive[k], 0.0f), 0.0f);
  p.double_sided = (m.flags & OXR_MAT_DOUBLE_SIDED) != 0;
  p.unlit = (m.flags & OXR_MAT_UNLIT) != 0;
  p.alpha_mode = m.alpha_mode;
  p.alpha_cutoff = m.alpha_cutoff;
  p.ior = fmaxr(fin(m.ior, 1.5f), 1.0f);
  p.specular = clampr(fin(m.specular, 1.0f), 0.0f, 1.0f);
  for (int k = 0; k < 3; ++k) p.specular_color[k] = fmaxr(fin(m.specular_color[k], 1.0f), 0.0f);
  p.transmission = clampr(fin(m.transmission, 0.0f), 0.0f, 1.0f);
  p.thickness = fmaxr(fin(m.thickness, 0.0f), 0.0f);
  for (int k = 0; k < 3; ++k) p.attenuation_color[k] = clampr(fin(m.attenuation_color[k], 1.0f), 0.0f, 1.0f);
  float d = m.attenuation_distance;
  p.attenuation_distance = (std::isfinite(d) && d > 0.0f) ? d : INF_F;
  p.clearcoat = clampr(fin(m.clearcoat, 0.0f), 0.0f, 1.0f);
  p.clearcoat_roughness = clampr(fin(m.clearcoat_roughness, 0.0f), 0.0f, 1.0f);
  for (int k = 0; k < 3; ++k) p.sheen_color[k] = clampr(fin(m.sheen_color[k], 0.0f), 0.0f, 1.0f);
  p.sheen_roughness = clampr(fin(m.sheen_roughness, 0.0f), 0.0f, 1.0f);
  p.iridescence = clampr(fin(m.iridescence, 0.0f), 0.0f, 1.0f);
  p.iridescence_ior = fmaxr(fin(m.iridescence_ior, 1.3f), 1.0f);
  p.iridescence_thickness = fmaxr(fin(m.iridescence_thickness, 400.0f), 0.0f);
  p.diffuse_transmission = clampr(fin(m.diffuse_transmission, 0.0f), 0.0f, 1.0f);
  for (int k = 0; k < 3; ++k)
    p.diffuse_transmission_color[k] = clampr(fin(m.diffuse_transmission_color[k], 1.0f), 0.0f, 1.0f);
  p.dispersion = fmaxr(fin(m.dispersion, 0.0f), 0.0f);
  p.anisotropy = clampr(fin(m.anisotropy, 0.0f), 0.0f, 1.0f);
  p.anisotropy_rotation = fin(m.anisotropy_rotation, 0.0f);
  return p;
}

struct PLight {  // PreparedLight (src/prepare.rs:693-714)
  uint32_t kind;
  f3 color;
  float intensity;
  f3 position, direction;
  bool has_range;
  float range;
  float inner, outer;
};

// PreparedLight::from_scene's sanitising (src/prepare.rs:866-919).
static PLight prepare_light(const oxr_light& l) {
  auto fin0 = [](float v) { return std::isfinite(v) ? fmaxr(v, 0.0f) : 0.0f; };
  PLight p;
  p.kind = l.kind;
  p.color = {fin0(l.color[0]), fin0(l.color[1]), fin0(l.color[2])};
  p.intensity = fin0(l.intensity);
  p.position = {l.position[0], l.position[1], l.position[2]};
  p.direction = {l.direction[0], l.direction[1], l.direction[2]};
  p.has_range = std::isfinite(l.range) && l.range > 0.0f;
  p.range = l.range;
  p.inner = 0.0f;
  p.outer = 0.0f;
  if (l.kind == OXR_LIGHT_SPOT) {
    float outer = std::isfinite(l.outer_cone_angle) ? clampr(l.outer_cone_angle, 1.0e-4f, FRAC_PI_2_F) : FRAC_PI_4_F;
    float inner = std::isfinite(l.inner_cone_angle) ? clampr(l.inner_cone_angle, 0.0f, outer) : 0.0f;
    p.inner = inner;
    p.outer = outer;
  }
  if (l.kind == OXR_LIGHT_DIRECTIONAL) p.has_range = false;
  return p;
}

struct LightSample {
  f3 l;
  float distance;
  f3 radiance;
};

// PreparedLight::sample (src/prepare.rs:921-963).
static bool light_sample(const PLight& L, f3 p, LightSample& out) {
  f3 base = {L.color[0] * L.intensity, L.color[1] * L.intensity, L.color[2] * L.intensity};
  if (L.kind == OXR_LIGHT_DIRECTIONAL) {
    out.l = neg(L.direction);
    out.distance = INF_F;
    out.radiance = base;
    return true;
  }
  f3 to = sub(L.position, p);
  float d2 = dot(to, to);
  if (std::isnan(d2) || d2 <= 1.0e-12f) return false;
  float d = std::sqrt(d2);
  f3 l = {to[0] / d, to[1] / d, to[2] / d};
  float att = 1.0f / d2;
  if (L.has_range) {
    float x = d / L.range;
    att *= clampr(1.0f - x * x * x * x, 0.0f, 1.0f);
  }
  if (L.kind == OXR_LIGHT_SPOT) {
    float cos_outer = std::cos(L.outer);
    float cos_inner = std::cos(L.inner);
    float sc = 1.0f / fmaxr(cos_inner - cos_outer, 0.001f);
    float offset = -cos_outer * sc;
    float cd = dot(L.direction, neg(l));
    float a = clampr(cd * sc + offset, 0.0f, 1.0f);
    att *= a * a;
  }
  if (att <= 0.0f) return false;
  out.l = l;
  out.distance = d;
  out.radiance = {base[0] * att, base[1] * att, base[2] * att};
  return true;
}

// Watertight ray/triangle test (Woop, Benthin, Wald 2013).
struct PRay {
  f3 o, d;
  int kx, ky, kz;
  float Sx, Sy, Sz;
  f3 inv;
};
static PRay prepare_ray(f3 o, f3 d) {
  PRay r;
  r.o = o;
  r.d = d;
  int kz = 0;
  if (std::fabs(d[1]) > std::fabs(d[kz])) kz = 1;
  if (std::fabs(d[2]) > std::fabs(d[kz])) kz = 2;
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  if (d[kz] < 0.0f) std::swap(kx, ky);
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  r.Sx = d[kx] / d[kz];
  r.Sy = d[ky] / d[kz];
  r.Sz = 1.0f / d[kz];
  for (int k = 0; k < 3; ++k) r.inv[k] = 1.0f / d[k];
  return r;
}

// Returns true and (t, b1, b2, front) on a hit with t in [t_min, t_max].
static bool intersect_tri(const PRay& r, const float* v9, float t_min, float t_max, float& t_out, float& b1,
                          float& b2, bool& front) {
  f3 A = {v9[0] - r.o[0], v9[1] - r.o[1], v9[2] - r.o[2]};
  f3 B = {v9[3] - r.o[0], v9[4] - r.o[1], v9[5] - r.o[2]};
  f3 C = {v9[6] - r.o[0], v9[7] - r.o[1], v9[8] - r.o[2]};
  float Ax = A[r.kx] - r.Sx * A[r.kz];
  float Ay = A[r.ky] - r.Sy * A[r.kz];
  float Bx = B[r.kx] - r.Sx * B[r.kz];
  float By = B[r.ky] - r.Sy * B[r.kz];
  float Cx = C[r.kx] - r.Sx * C[r.kz];
  float Cy = C[r.ky] - r.Sy * C[r.kz];
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = U + V + W;
  if (!(det != 0.0f)) return false;
  float Az = r.Sz * A[r.kz];
  float Bz = r.Sz * B[r.kz];
  float Cz = r.Sz * C[r.kz];
  float T = U * Az + V * Bz + W * Cz;
  float rcp = 1.0f / det;
  float t = T * rcp;
  if (!(t >= t_min && t <= t_max)) return false;
  t_out = t;
  b1 = V * rcp;
  b2 = W * rcp;
  front = det > 0.0f;
  return true;
}

struct Hit {
  float t;
  uint32_t g;
  f3 bary;
  bool front;
};

// Binary binned-SAH BVH over the soup (the oracle's own; Wald 2007).
struct BNode {
  float bmin[3], bmax[3];
  uint32_t left_or_first;  // inner: left child index; leaf: first slot
  uint32_t count;          // 0 = inner (right = left + 1)
};

// Equirectangular environment with luminance x sin(theta) importance
// sampling (EnvironmentMap, src/pathtrace.rs:1557-1741).
struct EnvMap {
  uint32_t w = 1, h = 1;
  std::vector<float> img;  // w*h*4, row 0 = +Y
  float intensity = 1.0f;
  std::vector<float> marginal, conditional, mass;
  bool uniform = true;

  // EnvironmentMap::new (:1575-1628).
  void build(const float* rgba, uint32_t w_, uint32_t h_, float inten) {
    w = std::max<uint32_t>(w_, 1);
    h = std::max<uint32_t>(h_, 1);
    bool valid = rgba && w_ > 0 && h_ > 0;
    img.assign(rgba ? rgba : nullptr, rgba ? rgba + (size_t)w * h * 4 : nullptr);
    std::vector<double> weights((size_t)w * h, 0.0);
    if (valid) {
      for (uint32_t y = 0; y < h; ++y) {
        double st = (double)std::sin(PI_F * ((float)y + 0.5f) / (float)h);
        for (uint32_t x = 0; x < w; ++x) {
          const float* px = &img[((size_t)y * w + x) * 4];
          float l = luminance({px[0], px[1], px[2]});
          weights[(size_t)y * w + x] = (std::isfinite(l) && l > 0.0f) ? (double)l * st : 0.0;
        }
      }
    }
    double total = 0.0;
    for (double v : weights) total += v;
    uniform = !(total > 0.0 && std::isfinite(total));
    if (uniform)
      for (size_t i = 0; i < weights.size(); ++i)
        weights[i] = (double)std::sin(PI_F * ((float)(i / w) + 0.5f) / (float)h);
    total = 0.0;
    for (double v : weights) total += v;
    marginal.assign(1, 0.0f);
    conditional.clear();
    double acc = 0.0;
    for (uint32_t y = 0; y < h; ++y) {
      double row = 0.0;
      for (uint32_t x = 0; x < w; ++x) row += weights[(size_t)y * w + x];
      double racc = 0.0;
      conditional.push_back(0.0f);
      for (uint32_t x = 0; x < w; ++x) {
        racc += weights[(size_t)y * w + x];
        conditional.push_back(row > 0.0 ? (float)(racc / row) : (float)(x + 1) / (float)w);
      }
      acc += row;
      marginal.push_back((float)(acc / total));
    }
    marginal.back() = 1.0f;
    mass.resize(weights.size());
    for (size_t i = 0; i < weights.size(); ++i) mass[i] = (float)(weights[i] / total);
    intensity = std::isfinite(inten) ? fmaxr(inten, 0.0f) : 0.0f;
  }

  void texel_of(f3 d, uint32_t& x, uint32_t& y, float& st) const {
    float u = 0.5f + std::atan2(d[0], -d[2]) / (2.0f * PI_F);
    float th = std::acos(clampr(d[1], -1.0f, 1.0f));
    float v = th / PI_F;
    float fx = u * (float)w, fy = v * (float)h;
    x = std::min<uint32_t>(fx > 0.0f ? (uint32_t)fx : 0u, w - 1);
    y = std::min<uint32_t>(fy > 0.0f ? (uint32_t)fy : 0u, h - 1);
    st = std::sin(th);
  }
  f3 radiance(f3 d) const {
    uint32_t x, y;
    float st;
    texel_of(d, x, y, st);
    if (uniform) return {0, 0, 0};
    const float* p = &img[((size_t)y * w + x) * 4];
    f3 c = {p[0], p[1], p[2]};
    if (!finite3(c)) return {0, 0, 0};
    return scale({fmaxr(c[0], 0.0f), fmaxr(c[1], 0.0f), fmaxr(c[2], 0.0f)}, intensity);
  }
  float pdf(f3 d) const {
    uint32_t x, y;
    float st;
    texel_of(d, x, y, st);
    if (st <= 1.0e-6f) return 0.0f;
    return mass[(size_t)y * w + x] * (float)(w * h) / (2.0f * PI_F * PI_F * st);
  }
  // EnvironmentMap::sample (:1713-1736).
  bool sample(float u0, float u1, f3& d, f3& le, float& pdf_out) const {
    auto ppoint = [](const float* c, size_t n, float u) {
      size_t lo = 0, hi = n;
      while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (c[mid] <= u) lo = mid + 1;
        else hi = mid;
      }
      return lo;
    };
    size_t y = std::min<size_t>(std::max<size_t>(ppoint(marginal.data(), h + 1, u0), 1), h) - 1;
    const float* row = &conditional[y * (w + 1)];
    size_t x = std::min<size_t>(std::max<size_t>(ppoint(row, w + 1, u1), 1), w) - 1;
    float my = (u0 - marginal[y]) / fmaxr(marginal[y + 1] - marginal[y], 1e-12f);
    float mx = (u1 - row[x]) / fmaxr(row[x + 1] - row[x], 1e-12f);
    float u = ((float)x + clampr(mx, 0.0f, 0.99999f)) / (float)w;
    float v = ((float)y + clampr(my, 0.0f, 0.99999f)) / (float)h;
    float phi = 2.0f * PI_F * (u - 0.5f);
    float th = PI_F * v;
    float st = std::sin(th), ct = std::cos(th);
    d = {st * std::sin(phi), ct, -st * std::cos(phi)};
    pdf_out = pdf(d);
    if (!(pdf_out > 0.0f) || !std::isfinite(pdf_out)) return false;
    le = radiance(d);
    return true;
  }
};

// ---------------------------------------------------------------------
// Textures (src/texture.rs:225-515): linear-light mip chains, the glTF
// sampler's filters and wrap modes, and the texture slots of
// TraceScene::material_oriented (src/trace.rs:517-605, 640-895).
// ---------------------------------------------------------------------
using f4 = std::array<float, 4>;

struct MipLevel {
  uint32_t w = 0, h = 0;
  std::vector<f4> t;
};

// wrap_index (src/texture.rs:280-289); i64 as the reference.
static int64_t rem_euclid(int64_t a, int64_t n) {
  int64_t r = a % n;
  return r < 0 ? r + n : r;
}
static int64_t wrap_index(int64_t i, int64_t n, uint32_t mode) {
  if (mode == OXR_WRAP_CLAMP_TO_EDGE) return std::min(std::max(i, (int64_t)0), n - 1);
  if (mode == OXR_WRAP_MIRRORED_REPEAT) {
    int64_t m = rem_euclid(i, 2 * n);
    return m >= n ? 2 * n - 1 - m : m;
  }
  return rem_euclid(i, n);
}
static f4 fetch(const MipLevel& l, int64_t x, int64_t y, uint32_t ws, uint32_t wt) {
  int64_t xi = wrap_index(x, (int64_t)l.w, ws), yi = wrap_index(y, (int64_t)l.h, wt);
  return l.t[(size_t)yi * l.w + (size_t)xi];
}
static f4 lerp4(const f4& a, const f4& b, float t) {
  return {a[0] + (b[0] - a[0]) * t, a[1] + (b[1] - a[1]) * t, a[2] + (b[2] - a[2]) * t, a[3] + (b[3] - a[3]) * t};
}

static float srgb_decode(float c) {  // hdr.rs srgb_to_linear
  if (c <= 0.04045f) return c / 12.92f;
  return std::pow((c + 0.055f) / 1.055f, 2.4f);
}

// build_mips (src/texture.rs:323-381): level 0 from the image (sRGB bytes
// decoded through the 256-entry table, or linear bytes / 255; float
// texels as given), then 2x2 box filters down to 1x1.
static std::vector<MipLevel> build_mips(const oxr_texture& tx, bool srgb) {
  static const std::array<float, 256> lut = [] {
    std::array<float, 256> t{};
    for (int i = 0; i < 256; ++i) t[i] = srgb_decode((float)i / 255.0f);
    return t;
  }();
  MipLevel l0;
  l0.w = tx.width;
  l0.h = tx.height;
  size_t n = (size_t)tx.width * tx.height;
  l0.t.resize(n);
  if (tx.format == OXR_TEXFMT_RGBA8) {
    const uint8_t* p = static_cast<const uint8_t*>(tx.pixels);
    for (size_t i = 0; i < n; ++i) {
      for (int c = 0; c < 3; ++c) l0.t[i][c] = srgb ? lut[p[4 * i + c]] : (float)p[4 * i + c] / 255.0f;
      l0.t[i][3] = (float)p[4 * i + 3] / 255.0f;
    }
  } else {
    const float* p = static_cast<const float*>(tx.pixels);
    for (size_t i = 0; i < n; ++i)
      for (int c = 0; c < 4; ++c) l0.t[i][c] = p[4 * i + c];
  }
  std::vector<MipLevel> levels;
  levels.push_back(std::move(l0));
  while (!(levels.back().w == 1 && levels.back().h == 1)) {
    const MipLevel& pv = levels.back();
    MipLevel nx;
    nx.w = std::max(pv.w / 2, 1u);
    nx.h = std::max(pv.h / 2, 1u);
    nx.t.resize((size_t)nx.w * nx.h);
    for (uint32_t y = 0; y < nx.h; ++y)
      for (uint32_t x = 0; x < nx.w; ++x) {
        size_t x0 = std::min(2 * x, pv.w - 1), x1 = std::min(2 * x + 1, pv.w - 1);
        size_t y0 = std::min(2 * y, pv.h - 1), y1 = std::min(2 * y + 1, pv.h - 1);
        size_t pw = pv.w;
        f4 acc = {0, 0, 0, 0};
        for (size_t idx : {y0 * pw + x0, y0 * pw + x1, y1 * pw + x0, y1 * pw + x1})
          for (int c = 0; c < 4; ++c) acc[c] += pv.t[idx][c];
        nx.t[(size_t)y * nx.w + x] = {acc[0] * 0.25f, acc[1] * 0.25f, acc[2] * 0.25f, acc[3] * 0.25f};
      }
    levels.push_back(std::move(nx));
  }
  return levels;
}

// filter_level (src/texture.rs:481-512).
static f4 filter_level(const MipLevel& l, float u, float v, bool linear, uint32_t ws, uint32_t wt) {
  u = std::isfinite(u) ? u : 0.0f;
  v = std::isfinite(v) ? v : 0.0f;
  u = clampr(u, -1.0e6f, 1.0e6f);
  v = clampr(v, -1.0e6f, 1.0e6f);
  float x = u * (float)l.w, y = v * (float)l.h;
  if (!linear) return fetch(l, (int64_t)std::floor(x), (int64_t)std::floor(y), ws, wt);
  float xf = x - 0.5f, yf = y - 0.5f;
  float x0 = std::floor(xf), y0 = std::floor(yf);
  float fx = xf - x0, fy = yf - y0;
  int64_t ix = (int64_t)x0, iy = (int64_t)y0;
  f4 a = fetch(l, ix, iy, ws, wt), b = fetch(l, ix + 1, iy, ws, wt);
  f4 c = fetch(l, ix, iy + 1, ws, wt), d = fetch(l, ix + 1, iy + 1, ws, wt);
  return lerp4(lerp4(a, b, fx), lerp4(c, d, fx), fy);
}

struct PTexture {  // PreparedTexture: image + sampler, both chains
  oxr_texture desc{};
  std::vector<MipLevel> chain[2];  // 0: sRGB, 1: linear
};

// The texel filter of a minification filter (the base filter of
// oxideav-mesh3d's MinFilter: X_MIPMAP_Y filters texels with X).
static bool min_base_linear(uint32_t m) {
  return m == OXR_MIN_LINEAR || m == OXR_MIN_LINEAR_MIP_NEAREST || m == OXR_MIN_LINEAR_MIP_LINEAR;
}

// PreparedTexture::sample_lod (src/texture.rs:417-447).
static f4 sample_lod(const PTexture& t, bool srgb, float u, float v, float lod) {
  const std::vector<MipLevel>& mips = t.chain[srgb ? 0 : 1];
  const oxr_texture& s = t.desc;
  if (std::isnan(lod) || lod <= 0.0f) return filter_level(mips[0], u, v, s.mag_filter == OXR_FILTER_LINEAR, s.wrap_s, s.wrap_t);
  uint32_t mn = s.min_filter;
  bool lin = min_base_linear(mn);
  float max_level = (float)(mips.size() - 1);
  if (mn == OXR_MIN_NEAREST || mn == OXR_MIN_LINEAR) return filter_level(mips[0], u, v, lin, s.wrap_s, s.wrap_t);
  if (mn == OXR_MIN_NEAREST_MIP_NEAREST || mn == OXR_MIN_LINEAR_MIP_NEAREST) {
    size_t l = (size_t)clampr(std::round(lod), 0.0f, max_level);
    return filter_level(mips[l], u, v, lin, s.wrap_s, s.wrap_t);
  }
  float l = clampr(lod, 0.0f, max_level);
  size_t l0 = (size_t)std::floor(l);
  size_t l1 = std::min(l0 + 1, mips.size() - 1);
  float f = l - (float)l0;
  f4 a = filter_level(mips[l0], u, v, lin, s.wrap_s, s.wrap_t);
  if (f <= 0.0f || l0 == l1) return a;
  f4 b = filter_level(mips[l1], u, v, lin, s.wrap_s, s.wrap_t);
  return lerp4(a, b, f);
}

// PreparedTexture::sample_grad's level of detail (src/texture.rs:452-470).
static float grad_lod(const PTexture& t, float dux, float dvx, float duy, float dvy) {
  float w = (float)t.desc.width, h = (float)t.desc.height;
  float lx = std::hypot(dux * w, dvx * h), ly = std::hypot(duy * w, dvy * h);
  float rho = fmaxr(lx, ly);
  return (std::isfinite(rho) && rho > 0.0f) ? std::log2(rho) : 0.0f;
}

// A material's texture reference (TextureBinding, src/prepare.rs:386-431)
// with its KHR_texture_transform as the 2x3 matrix T R S. The rotation
// follows KHR_texture_transform's definition (oxideav-mesh3d's
// TextureTransform::apply, not vendored: restated from the extension's
// specification; parity unpinned).
struct TBind {
  int32_t tex = -1;
  uint32_t uv_set = 0;
  bool xf = false;
  float m00 = 1, m01 = 0, m02 = 0, m10 = 0, m11 = 1, m12 = 0;
};
static TBind make_bind(const oxr_texture_binding& b, uint32_t ntex) {
  TBind t;
  if (b.texture < 0 || (uint32_t)b.texture >= ntex) return t;
  t.tex = b.texture;
  t.uv_set = b.uv_set;
  t.xf = b.has_transform != 0;
  if (t.xf) {
    float c = std::cos(b.rotation), s = std::sin(b.rotation);
    t.m00 = c * b.scale[0];
    t.m01 = s * b.scale[1];
    t.m02 = b.offset[0];
    t.m10 = -s * b.scale[0];
    t.m11 = c * b.scale[1];
    t.m12 = b.offset[1];
  }
  return t;
}
static void transform_uv(const TBind& b, float& u, float& v) {
  if (!b.xf) return;
  float nu = b.m00 * u + b.m01 * v + b.m02;
  float nv = b.m10 * u + b.m11 * v + b.m12;
  u = nu;
  v = nv;
}

// The slots of oxr_material_textures, in its order.
enum TexSlot {
  TS_BASE, TS_MR, TS_NORMAL, TS_EMISSIVE, TS_SPEC, TS_SPEC_COLOR, TS_TRANSMISSION, TS_THICKNESS, TS_CC, TS_CC_ROUGH,
  TS_CC_NORMAL, TS_SHEEN_COLOR, TS_SHEEN_ROUGH, TS_ANISO, TS_IRID, TS_IRID_THICK, TS_DT, TS_DT_COLOR, TS_COUNT
};
struct PMatTex {
  TBind slot[TS_COUNT];
  float normal_scale = 1.0f, cc_normal_scale = 1.0f, irid_min = 100.0f;
  bool any = false;
};
static PMatTex prepare_mat_tex(const oxr_material_textures& m, uint32_t ntex) {
  const oxr_texture_binding* b[TS_COUNT] = {
      &m.base_color, &m.metallic_roughness, &m.normal, &m.emissive, &m.specular, &m.specular_color,
      &m.transmission, &m.thickness, &m.clearcoat, &m.clearcoat_roughness, &m.clearcoat_normal, &m.sheen_color,
      &m.sheen_roughness, &m.anisotropy, &m.iridescence, &m.iridescence_thickness, &m.diffuse_transmission,
      &m.diffuse_transmission_color};
  PMatTex p;
  for (int k = 0; k < TS_COUNT; ++k) {
    p.slot[k] = make_bind(*b[k], ntex);
    p.any = p.any || p.slot[k].tex >= 0;
  }
  p.normal_scale = m.normal_scale;
  p.cc_normal_scale = m.clearcoat_normal_scale;
  p.irid_min = m.iridescence_thickness_min;
  return p;
}

struct Scene {
  uint32_t n = 0;
  std::vector<float> pos;        // 9 per tri
  std::vector<float> nrm;        // 9 per tri (or empty)
  std::vector<uint8_t> has_nrm;  // per tri
  std::vector<uint32_t> mat_id;
  std::vector<PMaterial> mats;
  std::vector<PLight> lights;
  // Textures and per-triangle attributes (empty: untextured).
  std::vector<PTexture> texs;
  std::vector<PMatTex> mtex;     // per material, or empty
  std::vector<uint32_t> tflags;  // OXR_TRI_* per triangle
  std::vector<float> uv[2];      // 6 per triangle
  std::vector<float> tan;        // 12 per triangle
  std::vector<float> col;        // 12 per triangle
  // EmissiveLights (src/pathtrace.rs:1436-1555).
  std::vector<uint32_t> l_tris;
  std::vector<float> l_pmf, l_cdf;
  std::vector<uint32_t> l_lookup;
  // BVH.
  std::vector<BNode> nodes;
  std::vector<uint32_t> slots;  // leaf slot -> global id
  float pad = 0.0f;
  std::shared_ptr<EnvMap> env;  // PathTracer::set_environment (null: the ambient sky)

  const float* tri(uint32_t g) const { return &pos[9ull * g]; }
  const PMaterial& mat(uint32_t g) const { return mats[mat_id[g]]; }
  const PMatTex* mat_tex(uint32_t g) const {
    if (mtex.empty()) return nullptr;
    const PMatTex& m = mtex[mat_id[g]];
    return m.any ? &m : nullptr;
  }
  bool has(uint32_t g, uint32_t bit) const { return !tflags.empty() && (tflags[g] & bit) != 0; }
};

static void build_emissive(Scene& s) {
  std::vector<double> power;
  s.l_lookup.assign(s.n, NO_LIGHT);
  s.l_tris.clear();
  std::vector<float> powf_;
  // Mean emission per material: factor x the emissive texture's 1x1 mip
  // (src/pathtrace.rs:1484-1500).
  std::vector<f3> mean_e(s.mats.size());
  for (size_t i = 0; i < s.mats.size(); ++i) {
    const PMaterial& m = s.mats[i];
    f3 e = m.unlit ? f3{0, 0, 0} : m.emissive;
    if (!m.unlit && !s.mtex.empty() && s.mtex[i].slot[TS_EMISSIVE].tex >= 0) {
      const f4& c = s.texs[s.mtex[i].slot[TS_EMISSIVE].tex].chain[0].back().t[0];
      e = {e[0] * c[0], e[1] * c[1], e[2] * c[2]};
    }
    mean_e[i] = e;
  }
  for (uint32_t g = 0; g < s.n; ++g) {
    const PMaterial& m = s.mat(g);
    f3 e = mean_e[s.mat_id[g]];
    float lum = luminance(e);
    if (std::isnan(lum) || lum <= 0.0f) continue;
    const float* v = s.tri(g);
    f3 a = {v[0], v[1], v[2]}, b = {v[3], v[4], v[5]}, c = {v[6], v[7], v[8]};
    f3 cr = cross(sub(b, a), sub(c, a));
    float area = 0.5f * std::sqrt(dot(cr, cr));
    if (std::isnan(area) || area <= 0.0f || !finitef(area)) continue;
    float sides = m.double_sided ? 2.0f : 1.0f;
    s.l_lookup[g] = (uint32_t)s.l_tris.size();
    s.l_tris.push_back(g);
    powf_.push_back(area * lum * sides);
  }
  double total = 0.0;
  for (float p : powf_) total += (double)p;
  s.l_pmf.clear();
  s.l_cdf.clear();
  double acc = 0.0;
  for (float pw : powf_) {
    s.l_pmf.push_back((float)((double)pw / total));
    acc += (double)pw / total;
    s.l_cdf.push_back((float)acc);
  }
  if (!s.l_cdf.empty()) s.l_cdf.back() = 1.0f;
}

static size_t light_pick(const Scene& s, float u) {
  // partition_point(|c| c <= u).min(len - 1)
  size_t lo = 0, hi = s.l_cdf.size();
  while (lo < hi) {
    size_t mid = (lo + hi) / 2;
    if (s.l_cdf[mid] <= u)
      lo = mid + 1;
    else
      hi = mid;
  }
  return std::min(lo, s.l_tris.size() - 1);
}

struct BuildItem {
  float bmin[3], bmax[3], c[3];
  uint32_t g;
};

static void build_bvh(Scene& s) {
  s.nodes.clear();
  s.slots.clear();
  if (s.n == 0) return;
  std::vector<BuildItem> items(s.n);
  float gmax = 0.0f;
  for (uint32_t g = 0; g < s.n; ++g) {
    const float* v = s.tri(g);
    BuildItem& it = items[g];
    for (int k = 0; k < 3; ++k) {
      it.bmin[k] = std::fmin(v[k], std::fmin(v[3 + k], v[6 + k]));
      it.bmax[k] = std::fmax(v[k], std::fmax(v[3 + k], v[6 + k]));
      it.c[k] = 0.5f * (it.bmin[k] + it.bmax[k]);
      if (std::isfinite(it.bmin[k])) gmax = std::fmax(gmax, std::fabs(it.bmin[k]));
      if (std::isfinite(it.bmax[k])) gmax = std::fmax(gmax, std::fabs(it.bmax[k]));
    }
    it.g = g;
  }
  // Absolute padding of every box: covers the rounding of the
  // triangle test relative to the scene's coordinate magnitude.
  s.pad = std::ldexp(std::fmax(gmax, 1.0f), -20);
  struct Task {
    uint32_t node, lo, hi;
  };
  const int NB = 16;
  // One node of the top-down build: bounds, the binned-SAH split (or a
  // median split when binning cannot separate), and either a leaf or two
  // child tasks pushed right then left.
  auto split = [&](Task t, std::vector<BNode>& nodes, std::vector<uint32_t>& slots, std::vector<Task>& out) {
    BNode nd;
    float cmin[3] = {INF_F, INF_F, INF_F}, cmax[3] = {-INF_F, -INF_F, -INF_F};
    for (int k = 0; k < 3; ++k) {
      nd.bmin[k] = INF_F;
      nd.bmax[k] = -INF_F;
    }
    for (uint32_t i = t.lo; i < t.hi; ++i) {
      for (int k = 0; k < 3; ++k) {
        nd.bmin[k] = std::fmin(nd.bmin[k], items[i].bmin[k]);
        nd.bmax[k] = std::fmax(nd.bmax[k], items[i].bmax[k]);
        cmin[k] = std::fmin(cmin[k], items[i].c[k]);
        cmax[k] = std::fmax(cmax[k], items[i].c[k]);
      }
    }
    for (int k = 0; k < 3; ++k) {
      nd.bmin[k] -= s.pad;
      nd.bmax[k] += s.pad;
    }
    uint32_t cnt = t.hi - t.lo;
    int best_axis = -1;
    int best_bin = 0;
    if (cnt > 4) {
      double best_cost = std::numeric_limits<double>::infinity();
      for (int axis = 0; axis < 3; ++axis) {
        float ext = cmax[axis] - cmin[axis];
        if (!(ext > 0.0f) || !std::isfinite(ext)) continue;
        struct Bin {
          float bmin[3], bmax[3];
          uint32_t n;
        } bins[NB];
        for (auto& b : bins) {
          for (int k = 0; k < 3; ++k) {
            b.bmin[k] = INF_F;
            b.bmax[k] = -INF_F;
          }
          b.n = 0;
        }
        for (uint32_t i = t.lo; i < t.hi; ++i) {
          int bi = (int)((items[i].c[axis] - cmin[axis]) / ext * NB);
          bi = std::min(std::max(bi, 0), NB - 1);
          Bin& b = bins[bi];
          b.n++;
          for (int k = 0; k < 3; ++k) {
            b.bmin[k] = std::fmin(b.bmin[k], items[i].bmin[k]);
            b.bmax[k] = std::fmax(b.bmax[k], items[i].bmax[k]);
          }
        }
        auto area = [](const float* mn, const float* mx) {
          double dx = std::fmax(mx[0] - mn[0], 0.0f), dy = std::fmax(mx[1] - mn[1], 0.0f),
                 dz = std::fmax(mx[2] - mn[2], 0.0f);
          return dx * dy + dy * dz + dz * dx;
        };
        double left_a[NB], right_a[NB];
        uint32_t left_n[NB], right_n[NB];
        float mn[3] = {INF_F, INF_F, INF_F}, mx[3] = {-INF_F, -INF_F, -INF_F};
        uint32_t acc = 0;
        for (int i = 0; i < NB; ++i) {
          for (int k = 0; k < 3; ++k) {
            mn[k] = std::fmin(mn[k], bins[i].bmin[k]);
            mx[k] = std::fmax(mx[k], bins[i].bmax[k]);
          }
          acc += bins[i].n;
          left_a[i] = acc ? area(mn, mx) : 0.0;
          left_n[i] = acc;
        }
        for (int k = 0; k < 3; ++k) {
          mn[k] = INF_F;
          mx[k] = -INF_F;
        }
        acc = 0;
        for (int i = NB - 1; i >= 0; --i) {
          for (int k = 0; k < 3; ++k) {
            mn[k] = std::fmin(mn[k], bins[i].bmin[k]);
            mx[k] = std::fmax(mx[k], bins[i].bmax[k]);
          }
          acc += bins[i].n;
          right_a[i] = acc ? area(mn, mx) : 0.0;
          right_n[i] = acc;
        }
        for (int i = 0; i < NB - 1; ++i) {
          if (left_n[i] == 0 || right_n[i + 1] == 0) continue;
          double cost = left_a[i] * left_n[i] + right_a[i + 1] * right_n[i + 1];
          if (cost < best_cost) {
            best_cost = cost;
            best_axis = axis;
            best_bin = i + 1;
          }
        }
      }
    }
    uint32_t mid = t.lo;
    if (best_axis >= 0) {
      float ext = cmax[best_axis] - cmin[best_axis];
      auto it = std::partition(items.begin() + t.lo, items.begin() + t.hi, [&](const BuildItem& b) {
        int bi = (int)((b.c[best_axis] - cmin[best_axis]) / ext * NB);
        bi = std::min(std::max(bi, 0), NB - 1);
        return bi < best_bin;
      });
      mid = (uint32_t)(it - items.begin());
    }
    if (cnt > 4 && (mid == t.lo || mid == t.hi)) {
      // Median split along the largest centroid extent.
      int axis = 0;
      for (int k = 1; k < 3; ++k)
        if (cmax[k] - cmin[k] > cmax[axis] - cmin[axis]) axis = k;
      mid = t.lo + cnt / 2;
      std::nth_element(items.begin() + t.lo, items.begin() + mid, items.begin() + t.hi,
                       [&](const BuildItem& a, const BuildItem& b) {
                         return a.c[axis] < b.c[axis] || (a.c[axis] == b.c[axis] && a.g < b.g);
                       });
    }
    if (cnt <= 4) {
      nd.left_or_first = (uint32_t)slots.size();
      nd.count = cnt;
      for (uint32_t i = t.lo; i < t.hi; ++i) slots.push_back(items[i].g);
      nodes[t.node] = nd;
    } else {
      uint32_t left = (uint32_t)nodes.size();
      nodes.push_back(BNode{});
      nodes.push_back(BNode{});
      nd.left_or_first = left;
      nd.count = 0;
      nodes[t.node] = nd;
      out.push_back({left + 1, mid, t.hi});
      out.push_back({left, t.lo, mid});
    }
  };
  // The top of the tree is built here until every open task holds at most
  // n / (8 threads) triangles; those subtrees are then built by worker
  // threads into their own arrays (they partition disjoint item ranges)
  // and appended in task order. The tree is the one a serial build makes;
  // only the node numbering differs.
  s.nodes.push_back(BNode{});
  const uint32_t nthreads = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
  const uint32_t grain = std::max<uint32_t>(4096, s.n / (8 * nthreads));
  std::vector<Task> stack{{0, 0, s.n}}, frontier;
  while (!stack.empty()) {
    Task t = stack.back();
    stack.pop_back();
    if (t.hi - t.lo > grain || nthreads == 1)
      split(t, s.nodes, s.slots, stack);
    else
      frontier.push_back(t);
  }
  struct Sub {
    std::vector<BNode> nodes;
    std::vector<uint32_t> slots;
  };
  std::vector<Sub> subs(frontier.size());
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (;;) {
      size_t k = next.fetch_add(1);
      if (k >= frontier.size()) break;
      Sub& sb = subs[k];
      sb.nodes.push_back(BNode{});
      std::vector<Task> st{{0, frontier[k].lo, frontier[k].hi}};
      while (!st.empty()) {
        Task t = st.back();
        st.pop_back();
        split(t, sb.nodes, sb.slots, st);
      }
    }
  };
  {
    std::vector<std::thread> ts;
    for (uint32_t i = 0; i < nthreads && i < frontier.size(); ++i) ts.emplace_back(work);
    for (auto& th : ts) th.join();
  }
  for (size_t k = 0; k < frontier.size(); ++k) {
    const Sub& sb = subs[k];
    const uint32_t base = (uint32_t)s.nodes.size(), slot0 = (uint32_t)s.slots.size();
    // Local node i > 0 lands at base + i - 1; the local root replaces the
    // frontier task's placeholder.
    auto fix = [&](BNode nd) {
      nd.left_or_first = nd.count ? nd.left_or_first + slot0 : base + nd.left_or_first - 1;
      return nd;
    };
    s.nodes[frontier[k].node] = fix(sb.nodes[0]);
    for (size_t i = 1; i < sb.nodes.size(); ++i) s.nodes.push_back(fix(sb.nodes[i]));
    s.slots.insert(s.slots.end(), sb.slots.begin(), sb.slots.end());
  }
}

// Conservative slab test: per-axis near/far picked by the direction's
// sign (a NaN from 0·inf leaves the interval unconstrained), far scaled up
// (Ize 2013's robust traversal, with extra margin).
static inline bool slab(const PRay& r, const BNode& nd, float t_min, float t_max, float& tnear_out) {
  float tnear = t_min, tfar = t_max;
  for (int k = 0; k < 3; ++k) {
    float t0 = (nd.bmin[k] - r.o[k]) * r.inv[k];
    float t1 = (nd.bmax[k] - r.o[k]) * r.inv[k];
    bool negd = std::signbit(r.inv[k]);
    float tn = negd ? t1 : t0;
    float tf = negd ? t0 : t1;
    tnear = std::fmax(tnear, tn);
    tfar = std::fmin(tfar, tf);
  }
  tnear_out = tnear;
  return tnear <= tfar * (1.0f + 1.0f / 65536.0f);
}

template <class Filter>
static bool closest_hit(const Scene& s, f3 o, f3 d, float t_min, float t_max, Filter&& accept, Hit& out,
                        bool brute = false) {
  if (s.n == 0) return false;
  PRay r = prepare_ray(o, d);
  bool found = false;
  float best_t = t_max;
  uint32_t best_g = 0;
  float bb1 = 0, bb2 = 0;
  bool bfront = false;
  auto test = [&](uint32_t g) {
    float t, b1, b2;
    bool front;
    if (!intersect_tri(r, s.tri(g), t_min, best_t, t, b1, b2, front)) return;
    if (found && !(t < best_t || (t == best_t && g < best_g))) return;
    Hit cand{t, g, {1.0f - b1 - b2, b1, b2}, front};
    if (!accept(cand)) return;
    found = true;
    best_t = t;
    best_g = g;
    bb1 = b1;
    bb2 = b2;
    bfront = front;
  };
  if (brute) {
    for (uint32_t g = 0; g < s.n; ++g) test(g);
  } else {
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const BNode& nd = s.nodes[stack[--sp]];
      float tn;
      if (!slab(r, nd, t_min, best_t, tn)) continue;
      if (nd.count > 0) {
        for (uint32_t i = 0; i < nd.count; ++i) test(s.slots[nd.left_or_first + i]);
      } else {
        stack[sp++] = nd.left_or_first + 1;
        stack[sp++] = nd.left_or_first;
      }
    }
  }
  if (!found) return false;
  out = Hit{best_t, best_g, {1.0f - bb1 - bb2, bb1, bb2}, bfront};
  return true;
}

template <class Filter>
static bool occluded(const Scene& s, f3 o, f3 d, float t_min, float t_max, Filter&& accept, bool brute = false) {
  if (s.n == 0) return false;
  PRay r = prepare_ray(o, d);
  auto test = [&](uint32_t g) {
    float t, b1, b2;
    bool front;
    if (!intersect_tri(r, s.tri(g), t_min, t_max, t, b1, b2, front)) return false;
    Hit cand{t, g, {1.0f - b1 - b2, b1, b2}, front};
    return accept(cand);
  };
  if (brute) {
    for (uint32_t g = 0; g < s.n; ++g)
      if (test(g)) return true;
    return false;
  }
  uint32_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const BNode& nd = s.nodes[stack[--sp]];
    float tn;
    if (!slab(r, nd, t_min, t_max, tn)) continue;
    if (nd.count > 0) {
      for (uint32_t i = 0; i < nd.count; ++i)
        if (test(s.slots[nd.left_or_first + i])) return true;
    } else {
      stack[sp++] = nd.left_or_first + 1;
      stack[sp++] = nd.left_or_first;
    }
  }
  return false;
}

struct Surface {
  f3 position, geometric_normal, shading_normal;
  float area;
};

// TraceScene::surface (src/trace.rs:459-488).
static Surface surface(const Scene& s, uint32_t g, f3 b) {
  const float* p = s.tri(g);
  Surface sf;
  sf.position = interp3(p, b);
  f3 p0 = {p[0], p[1], p[2]}, p1 = {p[3], p[4], p[5]}, p2 = {p[6], p[7], p[8]};
  f3 cr = cross(sub(p1, p0), sub(p2, p0));
  float len = std::sqrt(dot(cr, cr));
  sf.geometric_normal = len > 0.0f ? f3{cr[0] / len, cr[1] / len, cr[2] / len} : f3{0.0f, 0.0f, 1.0f};
  if (!s.nrm.empty() && s.has_nrm[g]) {
    f3 n = normalise(interp3(&s.nrm[9ull * g], b));
    sf.shading_normal = dot(n, n) > 0.5f ? n : sf.geometric_normal;
  } else {
    sf.shading_normal = sf.geometric_normal;
  }
  sf.area = 0.5f * len;
  return sf;
}

// Level of detail of a texture tap (TexLod): the ray cone at the first
// hit, else level 0 (src/pathtrace.rs:2083-2092).
struct TexLod {
  bool cone = false;
  float width = 0.0f;
  f3 dir = {0, 0, 0};
};

static bool tri_uvs(const Scene& s, uint32_t g, uint32_t set, const float*& uv) {
  if (set > 1 || !s.has(g, set ? OXR_TRI_HAS_UV1 : OXR_TRI_HAS_UV0)) return false;
  uv = &s.uv[set][6ull * g];
  return true;
}
static f4 interp4v(const float* v12, f3 b) {
  f4 o;
  for (int k = 0; k < 4; ++k) o[k] = v12[k] * b[0] + v12[4 + k] * b[1] + v12[8 + k] * b[2];
  return o;
}

// TraceScene::lod (src/trace.rs:518-552): the ray-cone level for a binding.
static float cone_lod(const Scene& s, const TBind& b, uint32_t g, const float* uvs, const TexLod& L) {
  float uv[3][2];
  for (int k = 0; k < 3; ++k) {
    uv[k][0] = uvs[2 * k];
    uv[k][1] = uvs[2 * k + 1];
    transform_uv(b, uv[k][0], uv[k][1]);
  }
  const oxr_texture& d = s.texs[b.tex].desc;
  float ta = (float)d.width * (float)d.height *
             std::fabs((uv[1][0] - uv[0][0]) * (uv[2][1] - uv[0][1]) - (uv[2][0] - uv[0][0]) * (uv[1][1] - uv[0][1]));
  const float* p = s.tri(g);
  f3 p0 = {p[0], p[1], p[2]}, p1 = {p[3], p[4], p[5]}, p2 = {p[6], p[7], p[8]};
  f3 cr = cross(sub(p1, p0), sub(p2, p0));
  float pa = std::sqrt(dot(cr, cr));
  if (!(ta > 0.0f && pa > 0.0f && L.width > 0.0f)) return 0.0f;
  f3 n = {cr[0] / pa, cr[1] / pa, cr[2] / pa};
  float cos = fmaxr(std::fabs(dot(n, L.dir)), 1.0e-3f);
  float l = 0.5f * std::log2(ta / pa) + std::log2(L.width / cos);
  return std::isfinite(l) ? l : 0.0f;
}

// TraceScene::sample_at (src/trace.rs:554-590): the binding's texel value
// at the hit, or false when the texture or its UV set is missing.
static bool tex_tap(const Scene& s, const TBind& b, uint32_t g, f3 bary, const TexLod& L, bool srgb, f4& out) {
  const float* uvs;
  if (b.tex < 0 || !tri_uvs(s, g, b.uv_set, uvs)) return false;
  float lod = L.cone ? cone_lod(s, b, g, uvs, L) : 0.0f;
  float u = uvs[0] * bary[0] + uvs[2] * bary[1] + uvs[4] * bary[2];
  float v = uvs[1] * bary[0] + uvs[3] * bary[1] + uvs[5] * bary[2];
  transform_uv(b, u, v);
  out = sample_lod(s.texs[b.tex], srgb, u, v, lod);
  return true;
}

// TraceScene::tangent_frame (src/trace.rs:640-663).
static bool tangent_frame(const Scene& s, uint32_t g, f3 b, f3 n, f3& tt, f3& bt) {
  if (!s.has(g, OXR_TRI_HAS_TANGENTS)) return false;
  const float* tg12 = &s.tan[12ull * g];
  f4 tg = interp4v(tg12, b);
  float w = tg12[3] < 0.0f ? -1.0f : 1.0f;
  f3 t3 = {tg[0], tg[1], tg[2]};
  float d = dot(n, t3);
  tt = normalise({t3[0] - n[0] * d, t3[1] - n[1] * d, t3[2] - n[2] * d});
  if (dot(tt, tt) < 0.5f) return false;
  f3 c = cross(n, tt);
  bt = {c[0] * w, c[1] * w, c[2] * w};
  return true;
}

// TraceScene::perturb (src/trace.rs:665-687).
static f3 perturb(const Scene& s, uint32_t g, f3 b, f3 n, const f4& t, float scale) {
  f3 tt, bt;
  if (!tangent_frame(s, g, b, n, tt, bt)) return n;
  float ts[3] = {(t[0] * 2.0f - 1.0f) * scale, (t[1] * 2.0f - 1.0f) * scale, t[2] * 2.0f - 1.0f};
  f3 nn = normalise({tt[0] * ts[0] + bt[0] * ts[1] + n[0] * ts[2], tt[1] * ts[0] + bt[1] * ts[1] + n[1] * ts[2],
                     tt[2] * ts[0] + bt[2] * ts[1] + n[2] * ts[2]});
  return dot(nn, nn) > 0.5f ? nn : n;
}

// TraceScene::material_oriented (src/trace.rs:700-923); `flip` reverses
// the shading normal before the normal maps are applied.
static MaterialSample material_at(const Scene& s, uint32_t g, f3 bary, const Surface& sf, bool flip,
                                  const TexLod& L) {
  const PMaterial& m = s.mat(g);
  const PMatTex* mt = s.mat_tex(g);
  f3 sn = flip ? neg(sf.shading_normal) : sf.shading_normal;
  MaterialSample ms;
  f4 t;
  auto tap = [&](int slot, bool srgb) { return mt && tex_tap(s, mt->slot[slot], g, bary, L, srgb, t); };
  for (int k = 0; k < 4; ++k) ms.base_color[k] = m.base_color[k];
  if (tap(TS_BASE, true))
    for (int k = 0; k < 4; ++k) ms.base_color[k] *= t[k];
  if (s.has(g, OXR_TRI_HAS_COLORS)) {
    f4 v = interp4v(&s.col[12ull * g], bary);
    for (int k = 0; k < 4; ++k) ms.base_color[k] *= v[k];
  }
  float metallic = m.metallic, roughness = m.roughness;
  if (tap(TS_MR, false)) {
    roughness *= t[1];
    metallic *= t[2];
  }
  ms.metallic = clampr(metallic, 0.0f, 1.0f);
  ms.roughness = clampr(roughness, 0.0f, 1.0f);
  ms.normal = sn;
  if (tap(TS_NORMAL, false)) ms.normal = perturb(s, g, bary, sn, t, mt->normal_scale);
  ms.emissive = m.emissive;
  if (tap(TS_EMISSIVE, true))
    for (int k = 0; k < 3; ++k) ms.emissive[k] *= t[k];
  ms.unlit = m.unlit;
  ms.double_sided = m.double_sided;
  ms.ior = m.ior;
  ms.specular = m.specular;
  if (tap(TS_SPEC, false)) ms.specular *= t[3];
  ms.specular_color = m.specular_color;
  if (tap(TS_SPEC_COLOR, true))
    for (int k = 0; k < 3; ++k) ms.specular_color[k] *= t[k];
  ms.transmission = m.transmission;
  if (tap(TS_TRANSMISSION, false)) ms.transmission *= t[0];
  ms.thickness = m.thickness;
  if (tap(TS_THICKNESS, false)) ms.thickness *= t[1];
  ms.attenuation_color = m.attenuation_color;
  ms.attenuation_distance = m.attenuation_distance;
  ms.clearcoat = m.clearcoat;
  if (tap(TS_CC, false)) ms.clearcoat *= t[0];
  ms.clearcoat_roughness = m.clearcoat_roughness;
  if (tap(TS_CC_ROUGH, false)) ms.clearcoat_roughness *= t[1];
  ms.clearcoat_normal = sn;
  if (tap(TS_CC_NORMAL, false)) ms.clearcoat_normal = perturb(s, g, bary, sn, t, mt->cc_normal_scale);
  ms.sheen_color = m.sheen_color;
  if (tap(TS_SHEEN_COLOR, true))
    for (int k = 0; k < 3; ++k) ms.sheen_color[k] *= t[k];
  ms.sheen_roughness = m.sheen_roughness;
  if (tap(TS_SHEEN_ROUGH, false)) ms.sheen_roughness *= t[3];
  // KHR_materials_anisotropy (src/trace.rs:835-858): the direction needs
  // the tangent frame; without one the lobe stays isotropic.
  ms.anisotropy = 0.0f;
  ms.anisotropy_direction = {0, 0, 0};
  if (m.anisotropy > 0.0f) {
    float a = m.anisotropy;
    float dir[2] = {1.0f, 0.0f};
    if (tap(TS_ANISO, false)) {
      dir[0] = t[0] * 2.0f - 1.0f;
      dir[1] = t[1] * 2.0f - 1.0f;
      a *= clampr(t[2], 0.0f, 1.0f);
    }
    float sr = std::sin(m.anisotropy_rotation), cr = std::cos(m.anisotropy_rotation);
    float rot[2] = {cr * dir[0] - sr * dir[1], sr * dir[0] + cr * dir[1]};
    f3 tt, bt;
    if (tangent_frame(s, g, bary, sn, tt, bt)) {
      f3 d = normalise({tt[0] * rot[0] + bt[0] * rot[1], tt[1] * rot[0] + bt[1] * rot[1],
                        tt[2] * rot[0] + bt[2] * rot[1]});
      if (dot(d, d) > 0.5f) ms.anisotropy_direction = d;
    }
    ms.anisotropy = dot(ms.anisotropy_direction, ms.anisotropy_direction) < 0.5f ? 0.0f : a;
  }
  ms.iridescence = m.iridescence;
  if (tap(TS_IRID, false)) ms.iridescence *= t[0];
  ms.iridescence_ior = m.iridescence_ior;
  ms.iridescence_thickness = m.iridescence_thickness;
  if (tap(TS_IRID_THICK, false)) {
    float lo = mt->irid_min, hi = m.iridescence_thickness;
    ms.iridescence_thickness = fmaxr(lo + (hi - lo) * clampr(t[1], 0.0f, 1.0f), 0.0f);
  }
  ms.diffuse_transmission = m.diffuse_transmission;
  if (tap(TS_DT, false)) ms.diffuse_transmission *= t[3];
  ms.diffuse_transmission_color = m.diffuse_transmission_color;
  if (tap(TS_DT_COLOR, true))
    for (int k = 0; k < 3; ++k) ms.diffuse_transmission_color[k] *= t[k];
  ms.dispersion = m.dispersion;
  return ms;
}

// TraceScene::alpha (src/trace.rs:492-506): base alpha x the base
// texture's alpha at level 0 x the vertex colour's alpha.
static float tri_alpha(const Scene& s, uint32_t g, f3 bary) {
  float a = s.mat(g).base_color[3];
  const PMatTex* mt = s.mat_tex(g);
  f4 t;
  if (mt && tex_tap(s, mt->slot[TS_BASE], g, bary, TexLod{}, true, t)) a *= t[3];
  if (s.has(g, OXR_TRI_HAS_COLORS)) a *= interp4v(&s.col[12ull * g], bary)[3];
  return a;
}

// ---------------------------------------------------------------------
// Integrator (src/pathtrace.rs:1877-2330).
// ---------------------------------------------------------------------
static inline float power_heuristic(float a, float b) {
  float a2 = a * a, b2 = b * b;
  if (a2 + b2 > 0.0f && finitef(a2 + b2)) return a2 / (a2 + b2);
  if (std::isinf(a)) return 1.0f;
  return 0.0f;
}

struct Integrator {
  const Scene& ts;
  oxr_camera cam;
  uint32_t width, height;
  float ambient;
  oxr_pt_options pt;
  std::vector<std::array<float, 8>>* record = nullptr;  // closest-hit rays, when recording
  std::vector<std::array<float, 8>>* record_shadow = nullptr;  // occlusion rays (o, skip bits, d, t_max)

  f3 clampc(f3 c) const {
    float m = max3(c);
    if (pt.clamp > 0.0f && m > pt.clamp) return scale(c, pt.clamp / m);
    return c;
  }
  // PathTracer::refine (src/pathtrace.rs:2537-2540): spread = 2 half_h /
  // height, the same for both projections (no atan).
  float pixel_spread() const { return 2.0f * cam.half_h / (float)std::max(height, 1u); }
  // Integrator::p_env (src/pathtrace.rs:1915-1921).
  float p_env() const { return ts.env ? (ts.l_tris.empty() ? 1.0f : 0.5f) : 0.0f; }

  bool accept(const Hit& c, const Sampler& s, uint32_t ray, bool primary) const {
    const PMaterial& mat = ts.mat(c.g);
    if (primary && !c.front && !mat.double_sided) return false;
    switch (mat.alpha_mode) {
      case OXR_ALPHA_MASK:
        return tri_alpha(ts, c.g, c.bary) >= mat.alpha_cutoff;
      case OXR_ALPHA_BLEND:
        return s.coin(ray, c.g) < tri_alpha(ts, c.g, c.bary);
      default:
        return true;
    }
  }
  bool visible(f3 o, f3 d, float t_max, const Sampler& s, uint32_t ray, uint32_t skip) const {
    if (record_shadow) {
      float sk;
      std::memcpy(&sk, &skip, 4);
      record_shadow->push_back({o[0], o[1], o[2], sk, d[0], d[1], d[2], t_max});
    }
    return !occluded(ts, o, d, 0.0f, t_max, [&](const Hit& c) { return c.g != skip && accept(c, s, ray, false); });
  }
  // Integrator::emission (src/pathtrace.rs:1951-1966): the factor times
  // the emissive texture at level 0.
  f3 emission(uint32_t g, f3 bary) const {
    const PMaterial& m = ts.mat(g);
    if (m.unlit || is_black(m.emissive)) return {0, 0, 0};
    f3 e = m.emissive;
    const PMatTex* mt = ts.mat_tex(g);
    f4 t;
    if (mt && tex_tap(ts, mt->slot[TS_EMISSIVE], g, bary, TexLod{}, true, t)) e = {e[0] * t[0], e[1] * t[1], e[2] * t[2]};
    return e;
  }
  float tri_pdf(size_t li, uint32_t g, f3 p, f3 q) const {
    const float* v = ts.tri(g);
    f3 a = {v[0], v[1], v[2]}, b = {v[3], v[4], v[5]}, c = {v[6], v[7], v[8]};
    float sel = (1.0f - p_env()) * ts.l_pmf[li];
    float omega = triangle_solid_angle(p, a, b, c);
    if (omega >= MIN_SPHERICAL_SOLID_ANGLE) return sel / omega;
    f3 cr = cross(sub(b, a), sub(c, a));
    float len = std::sqrt(dot(cr, cr));
    f3 to = sub(q, p);
    float d2 = dot(to, to);
    if (len <= 0.0f || d2 <= 0.0f) return 0.0f;
    float cos = std::fabs(dot(cr, to) / (len * std::sqrt(d2)));
    if (cos <= 0.0f) return 0.0f;
    return sel * d2 / (0.5f * len * cos);
  }
  float light_pdf(uint32_t g, f3 p, f3 q) const {
    uint32_t li = g < ts.l_lookup.size() ? ts.l_lookup[g] : NO_LIGHT;
    if (li == NO_LIGHT) return 0.0f;
    return tri_pdf(li, g, p, q);
  }

  void primary_ray(float px, float py, f3& o, f3& d) const {
    float ndc_x = (px / fmaxr((float)width, 1.0f)) * 2.0f - 1.0f;
    float ndc_y = 1.0f - (py / fmaxr((float)height, 1.0f)) * 2.0f;
    f3 side = {cam.side[0], cam.side[1], cam.side[2]};
    f3 up = {cam.up[0], cam.up[1], cam.up[2]};
    f3 fwd = {cam.forward[0], cam.forward[1], cam.forward[2]};
    f3 eye = {cam.eye[0], cam.eye[1], cam.eye[2]};
    if (cam.projection == OXR_PROJ_ORTHOGRAPHIC) {
      o = add(eye, add(scale(side, ndc_x * cam.half_w), scale(up, ndc_y * cam.half_h)));
      d = fwd;
    } else {
      d = normalise(add(add(scale(side, ndc_x * cam.half_w), scale(up, ndc_y * cam.half_h)), fwd));
      o = eye;
    }
  }

  // Integrator::sample (src/pathtrace.rs:2004-2329). Returns false for
  // an uncovered sample.
  bool sample(uint32_t x, uint32_t y, uint32_t index, f3& out) const {
    Sampler s(pt.seed, x, y, index);
    auto p0 = s.pattern(0);
    f3 o, d;
    primary_ray((float)x + p0[0], (float)y + p0[1], o, d);
    f3 beta = {1, 1, 1};
    f3 radiance = {0, 0, 0};
    float prev_pdf = 0.0f;
    f3 prev_pos = {0, 0, 0};
    bool has_medium = false;
    f3 sigma = {0, 0, 0};
    int hero = -1;
    uint32_t strategy = pt.strategy;
    uint32_t k = 0;
    for (;;) {
      uint32_t ray_id = k << 16;
      if (record) record->push_back({o[0], o[1], o[2], (float)k, d[0], d[1], d[2], INF_F});
      Hit hit;
      bool got = closest_hit(ts, o, d, 0.0f, INF_F, [&](const Hit& c) { return accept(c, s, ray_id, k == 0); }, hit);
      if (has_medium) {
        if (!got) break;
        for (int c = 0; c < 3; ++c) beta[c] *= std::exp(-sigma[c] * hit.t);
      }
      if (!got) {
        if (k == 0) return false;
        f3 le = {ambient, ambient, ambient};
        float w = 1.0f;
        if (ts.env) {  // src/pathtrace.rs:2038-2050
          le = ts.env->radiance(d);
          w = strategy == OXR_STRATEGY_MIS ? power_heuristic(prev_pdf, p_env() * ts.env->pdf(d))
              : strategy == OXR_STRATEGY_LIGHT_ONLY ? 0.0f : 1.0f;
        }
        if (w > 0.0f) radiance = add(radiance, clampc(scale(mul(beta, le), w)));
        break;
      }
      Surface surf = surface(ts, hit.g, hit.bary);
      const PMaterial& pmat = ts.mat(hit.g);
      if ((hit.front || pmat.double_sided) && !pmat.unlit && !is_black(pmat.emissive)) {
        f3 le = emission(hit.g, hit.bary);
        if (!is_black(le)) {
          if (k == 0) {
            radiance = add(radiance, mul(beta, le));
          } else {
            float pl = light_pdf(hit.g, prev_pos, surf.position);
            float w;
            if (pl <= 0.0f)
              w = 1.0f;
            else if (strategy == OXR_STRATEGY_MIS)
              w = power_heuristic(prev_pdf, pl);
            else if (strategy == OXR_STRATEGY_LIGHT_ONLY)
              w = 0.0f;
            else
              w = 1.0f;
            if (w > 0.0f) radiance = add(radiance, clampc(scale(mul(beta, le), w)));
          }
        }
      }
      bool flip = !hit.front && pmat.double_sided;
      TexLod lod;
      if (k == 0) {  // src/pathtrace.rs:2083-2092
        lod.cone = true;
        lod.width = cam.projection == OXR_PROJ_ORTHOGRAPHIC ? pixel_spread() : pixel_spread() * hit.t;
        lod.dir = d;
      }
      MaterialSample mat = material_at(ts, hit.g, hit.bary, surf, flip, lod);
      if (flip) {
        mat.normal = neg(mat.normal);
        mat.clearcoat_normal = neg(mat.clearcoat_normal);
      }
      if (mat.unlit) {
        f3 c = {mat.base_color[0], mat.base_color[1], mat.base_color[2]};
        f3 contrib = mul(beta, c);
        radiance = add(radiance, k == 0 ? contrib : clampc(contrib));
        break;
      }
      if (k >= pt.max_bounces) break;
      if (mat.dispersion > 0.0f && mat.thickness > 0.0f && mat.transmission > 0.0f) {
        int c;
        if (hero >= 0) {
          c = hero;
        } else {
          size_t ci = sat_usize(s.coin(ray_id | 0x7fffu, 0xffffffffu) * 3.0f);
          c = (int)std::min<size_t>(ci, 2);
          for (int i = 0; i < 3; ++i) beta[i] = (i == c) ? beta[i] * 3.0f : 0.0f;
          hero = c;
        }
        mat.ior = dispersion_iors(mat.ior, mat.dispersion)[c];
      }
      f3 v = neg(d);
      Bsdf bsdf(mat, v, surf.geometric_normal, hit.front);
      if (bsdf.black()) break;
      f3 p = surf.position;
      f3 ng = bsdf.ng;
      auto spawn = [&](f3 dir) { return dot(dir, ng) >= 0.0f ? offset_ray_origin(p, ng) : offset_ray_origin(p, neg(ng)); };

      // NEE: punctual lights.
      for (size_t j = 0; j < ts.lights.size(); ++j) {
        LightSample ls;
        if (!light_sample(ts.lights[j], p, ls)) continue;
        f3 f;
        float pdf_unused;
        bsdf.eval(ls.l, f, pdf_unused);
        if (is_black(f)) continue;
        f3 o2 = spawn(ls.l);
        float tmax = finitef(ls.distance) ? ls.distance * (1.0f - 1.0e-4f) : INF_F;
        if (visible(o2, ls.l, tmax, s, ray_id | 0x8000u | ((uint32_t)j & 0x7fffu), 0xffffffffu))
          radiance = add(radiance, clampc(mul(mul(beta, f), ls.radiance)));
      }

      // NEE: area set (emissive triangles / environment, src/pathtrace.rs:2173-2279).
      auto pl_u = s.pattern(2 + 2 * k);
      if (strategy != OXR_STRATEGY_BSDF_ONLY && (ts.env || !ts.l_tris.empty())) {
        float pe = p_env();
        uint32_t shadow_id = ray_id | 0xffffu;
        if (pl_u[2] < pe) {
          f3 l, le;
          float epdf;
          if (ts.env->sample(pl_u[0], pl_u[1], l, le, epdf)) {
            float pl = pe * epdf;
            f3 f;
            float pb;
            bsdf.eval(l, f, pb);
            if (!is_black(f) && !is_black(le) && pl > 0.0f) {
              float w = strategy == OXR_STRATEGY_MIS ? power_heuristic(pl, pb) : 1.0f;
              if (visible(spawn(l), l, INF_F, s, shadow_id, 0xffffffffu)) {
                f3 cc = scale(mul(mul(beta, f), le), w / pl);
                radiance = add(radiance, clampc(cc));
              }
            }
          }
        } else if (!ts.l_tris.empty()) {
          float us = pe > 0.0f ? clampr((pl_u[2] - pe) / (1.0f - pe), 0.0f, 0.999999f) : pl_u[2];
          size_t li = light_pick(ts, us);
          uint32_t g = ts.l_tris[li];
          const float* tv = ts.tri(g);
          f3 a = {tv[0], tv[1], tv[2]}, b = {tv[3], tv[4], tv[5]}, c = {tv[6], tv[7], tv[8]};
          bool have_target = false;
          f3 q;
          if (triangle_solid_angle(p, a, b, c) >= MIN_SPHERICAL_SOLID_ANGLE) {
            f3 l;
            if (sample_spherical_triangle(p, a, b, c, pl_u[0], pl_u[1], l)) {
              f3 n = cross(sub(b, a), sub(c, a));
              float dn = dot(l, n);
              if (dn != 0.0f) {
                float t = dot(sub(a, p), n) / dn;
                if (t > 0.0f && finitef(t)) {
                  q = add(p, scale(l, t));
                  have_target = true;
                }
              }
            }
          } else {
            float su = std::sqrt(pl_u[0]);
            float w[3] = {1.0f - su, su * (1.0f - pl_u[1]), su * pl_u[1]};
            q = {a[0] * w[0] + b[0] * w[1] + c[0] * w[2], a[1] * w[0] + b[1] * w[1] + c[1] * w[2],
                 a[2] * w[0] + b[2] * w[1] + c[2] * w[2]};
            have_target = true;
          }
          f3 bary;
          if (have_target && barycentric_of(q, a, b, c, bary)) {
            bary = {clampr(bary[0], 0.0f, 1.0f), clampr(bary[1], 0.0f, 1.0f), clampr(bary[2], 0.0f, 1.0f)};
            f3 to = sub(q, p);
            float dist2 = dot(to, to);
            Surface lsurf = surface(ts, g, bary);
            const PMaterial& lmat = ts.mat(g);
            if (dist2 > 1.0e-12f) {
              float dist = std::sqrt(dist2);
              f3 l = scale(to, 1.0f / dist);
              float cos_l = -dot(lsurf.geometric_normal, l);
              float pl = tri_pdf(li, g, p, q);
              if ((cos_l > 0.0f || (lmat.double_sided && cos_l < 0.0f)) && pl > 0.0f && finitef(pl)) {
                f3 f;
                float pb;
                bsdf.eval(l, f, pb);
                f3 le = emission(g, bary);
                if (!is_black(f) && !is_black(le)) {
                  float w = strategy == OXR_STRATEGY_MIS ? power_heuristic(pl, pb) : 1.0f;
                  f3 o2 = spawn(l);
                  float tmax = fmaxr(dist * (1.0f - 1.0e-4f), 0.0f);
                  if (visible(o2, l, tmax, s, shadow_id, g)) {
                    f3 cc = scale(mul(mul(beta, f), le), w / pl);
                    radiance = add(radiance, clampc(cc));
                  }
                }
              }
            }
          }
        }
      }

      // BSDF sampling.
      auto pb_u = s.pattern(1 + 2 * k);
      float uu[3] = {pb_u[0], pb_u[1], pb_u[2]};
      f3 l, f;
      float pdf;
      if (!bsdf.sample(uu, l, f, pdf)) break;
      beta = mul(beta, scale(f, 1.0f / pdf));
      if (!finite3(beta)) break;
      prev_pdf = pdf;
      prev_pos = p;
      if (dot(l, ng) < 0.0f && mat.thickness > 0.0f && mat.transmission > 0.0f) {
        if (hit.front) {
          float dist = mat.attenuation_distance;
          if (finitef(dist)) {
            f3 cc = mat.attenuation_color;
            sigma = {-std::log(fmaxr(cc[0], 1.0e-6f)) / dist, -std::log(fmaxr(cc[1], 1.0e-6f)) / dist,
                     -std::log(fmaxr(cc[2], 1.0e-6f)) / dist};
            has_medium = true;
          } else {
            has_medium = false;
          }
        } else {
          has_medium = false;
        }
      }
      if (k + 1 >= pt.rr_start) {
        float q = clampr(max3(beta), 0.05f, 1.0f);
        if (q < 1.0f) {
          if (pb_u[3] >= q) break;
          beta = scale(beta, 1.0f / q);
        }
      }
      o = spawn(l);
      d = l;
      k += 1;
    }
    out = finite3(radiance) ? radiance : f3{0, 0, 0};
    return true;
  }
};

// ---------------------------------------------------------------------
// Resolve (src/pathtrace.rs:2593-2734, src/hdr.rs:44-129).
// ---------------------------------------------------------------------
static float srgb_to_linear(float c) {
  if (c <= 0.04045f) return c / 12.92f;
  return std::pow((c + 0.055f) / 1.055f, 2.4f);
}
static float sanitize(float v) { return std::isnan(v) ? 0.0f : v; }
static float aces(float x) {
  x = x * 0.6f;
  float v = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
  return clampr(v, 0.0f, 1.0f);
}
static f3 tone_map(uint32_t tm, f3 rgb) {
  rgb = {fmaxr(sanitize(rgb[0]), 0.0f), fmaxr(sanitize(rgb[1]), 0.0f), fmaxr(sanitize(rgb[2]), 0.0f)};
  if (tm == OXR_TONEMAP_REINHARD) {
    float l = 0.2126f * rgb[0] + 0.7152f * rgb[1] + 0.0722f * rgb[2];
    if (l <= 0.0f) return {0, 0, 0};
    float s = 1.0f / (1.0f + l);
    return {fminr(rgb[0] * s, 1.0f), fminr(rgb[1] * s, 1.0f), fminr(rgb[2] * s, 1.0f)};
  }
  if (tm == OXR_TONEMAP_ACES) return {aces(rgb[0]), aces(rgb[1]), aces(rgb[2])};
  return {fminr(rgb[0], 1.0f), fminr(rgb[1], 1.0f), fminr(rgb[2], 1.0f)};
}
static uint8_t linear_to_srgb_byte(float c) {
  static float th[255];
  static std::once_flag once;
  std::call_once(once, [] {
    for (int k = 0; k < 255; ++k) th[k] = srgb_to_linear(((float)k + 0.5f) / 255.0f);
  });
  if (std::isnan(c)) return 0;
  int lo = 0, hi = 255;
  while (lo < hi) {
    int mid = (lo + hi) / 2;
    if (th[mid] <= c)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (uint8_t)lo;
}

}  // namespace oc

// =====================================================================
// C ABI of the oracle (loaded by tests/ and bench.py through ctypes).
// =====================================================================
using namespace oc;

extern "C" {

struct oxrcpu_scene {
  Scene s;
};

oxrcpu_scene* oxrcpu_scene_create(const oxr_scene_desc* d) {
  if (!d) return nullptr;
  auto* h = new oxrcpu_scene();
  Scene& s = h->s;
  s.n = d->num_tris;
  s.pos.assign(d->positions, d->positions + 9ull * s.n);
  s.has_nrm.assign(s.n, 0);
  if (d->normals) {
    s.nrm.assign(d->normals, d->normals + 9ull * s.n);
    for (uint32_t g = 0; g < s.n; ++g)
      s.has_nrm[g] = d->tri_flags ? ((d->tri_flags[g] & OXR_TRI_HAS_NORMALS) != 0) : 1;
  }
  s.mat_id.assign(d->material_ids, d->material_ids + s.n);
  for (uint32_t i = 0; i < d->num_materials; ++i) s.mats.push_back(prepare_material(d->materials[i]));
  for (uint32_t i = 0; i < d->num_lights; ++i) s.lights.push_back(prepare_light(d->lights[i]));
  // Textures: both chains of every texture (TextureData builds them lazily).
  for (uint32_t i = 0; i < d->num_textures; ++i) {
    PTexture t;
    t.desc = d->textures[i];
    if (t.desc.width == 0 || t.desc.height == 0 || !t.desc.pixels) {
      delete h;
      return nullptr;
    }
    t.chain[0] = build_mips(t.desc, true);
    t.chain[1] = build_mips(t.desc, false);
    t.desc.pixels = nullptr;
    s.texs.push_back(std::move(t));
  }
  if (d->material_textures)
    for (uint32_t i = 0; i < d->num_materials; ++i)
      s.mtex.push_back(prepare_mat_tex(d->material_textures[i], d->num_textures));
  const float* attr[4] = {d->uv0, d->uv1, d->tangents, d->colors};
  const uint32_t bits[4] = {OXR_TRI_HAS_UV0, OXR_TRI_HAS_UV1, OXR_TRI_HAS_TANGENTS, OXR_TRI_HAS_COLORS};
  const size_t width[4] = {6, 6, 12, 12};
  std::vector<float>* dst[4] = {&s.uv[0], &s.uv[1], &s.tan, &s.col};
  bool any_attr = false;
  for (int k = 0; k < 4; ++k)
    if (attr[k]) {
      dst[k]->assign(attr[k], attr[k] + width[k] * s.n);
      any_attr = true;
    }
  if (any_attr) {
    s.tflags.assign(s.n, 0u);
    for (uint32_t g = 0; g < s.n; ++g)
      for (int k = 0; k < 4; ++k)
        if (attr[k] && (!d->tri_flags || (d->tri_flags[g] & bits[k]))) s.tflags[g] |= bits[k];
  }
  build_emissive(s);
  build_bvh(s);
  return h;
}

void oxrcpu_scene_destroy(oxrcpu_scene* h) { delete h; }

// PathTracer::set_environment: rgba = w*h*4 floats (row 0 = +Y), or null
// to remove it.
void oxrcpu_set_environment(oxrcpu_scene* h, const float* rgba, uint32_t w, uint32_t hgt, float intensity) {
  if (!rgba) {
    h->s.env.reset();
    return;
  }
  auto e = std::make_shared<EnvMap>();
  e->build(rgba, w, hgt, intensity);
  h->s.env = e;
}

// EnvironmentMap::{pdf, radiance, sample} of the scene's map.
void oxrcpu_env_eval(oxrcpu_scene* h, uint32_t n, const float* dirs, float* pdf, float* le) {
  for (uint32_t i = 0; i < n; ++i) {
    f3 d = {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
    pdf[i] = h->s.env->pdf(d);
    f3 r = h->s.env->radiance(d);
    for (int c = 0; c < 3; ++c) le[3 * i + c] = r[c];
  }
}
void oxrcpu_env_sample(oxrcpu_scene* h, uint32_t n, const float* u, float* dirs, float* pdf, uint8_t* ok) {
  for (uint32_t i = 0; i < n; ++i) {
    f3 d, le;
    float p = 0.0f;
    ok[i] = h->s.env->sample(u[2 * i], u[2 * i + 1], d, le, p) ? 1 : 0;
    for (int c = 0; c < 3; ++c) dirs[3 * i + c] = d[c];
    pdf[i] = p;
  }
}
// The map's CDF tables (marginal h+1, conditional h*(w+1), mass w*h).
void oxrcpu_env_tables(oxrcpu_scene* h, float* marginal, float* conditional, float* mass, uint8_t* uniform) {
  const EnvMap& e = *h->s.env;
  std::copy(e.marginal.begin(), e.marginal.end(), marginal);
  std::copy(e.conditional.begin(), e.conditional.end(), conditional);
  std::copy(e.mass.begin(), e.mass.end(), mass);
  *uniform = e.uniform;
}

uint32_t oxrcpu_num_lights(oxrcpu_scene* h) { return (uint32_t)h->s.l_tris.size(); }

void oxrcpu_emissive_table(oxrcpu_scene* h, uint32_t* tris, float* pmf, float* cdf, uint32_t* lookup) {
  const Scene& s = h->s;
  for (size_t i = 0; i < s.l_tris.size(); ++i) {
    tris[i] = s.l_tris[i];
    pmf[i] = s.l_pmf[i];
    cdf[i] = s.l_cdf[i];
  }
  for (uint32_t g = 0; g < s.n; ++g) lookup[g] = s.l_lookup[g];
}

// Band-parallel refine (src/pathtrace.rs:2557-2589): rows [y0, y1) of a
// W x H frame, samples [first, first + n). accum is W*(y1-y0)*4 doubles
// (added to, so refines compose), sum_sq W*(y1-y0)*3 doubles or NULL.
int oxrcpu_render(oxrcpu_scene* h, const oxr_camera* cam, const oxr_pt_options* opts, uint32_t first, uint32_t n,
                  uint32_t y0, uint32_t y1, int threads, double* accum, double* sum_sq) {
  if (!h || !cam || !opts || !accum) return -1;
  uint32_t W = std::max<uint32_t>(opts->width, 1), H = std::max<uint32_t>(opts->height, 1);
  if (y1 > H || y0 >= y1) return -1;
  Integrator integ{h->s, *cam, W, H, std::isfinite(opts->ambient) ? fmaxr(opts->ambient, 0.0f) : 0.0f, *opts};
  const uint32_t BAND_ROWS = 4;
  uint32_t rows = y1 - y0;
  uint32_t nbands = (rows + BAND_ROWS - 1) / BAND_ROWS;
  std::atomic<uint32_t> next{0};
  auto work = [&]() {
    for (;;) {
      uint32_t band = next.fetch_add(1);
      if (band >= nbands) break;
      uint32_t r0 = band * BAND_ROWS, r1 = std::min(rows, r0 + BAND_ROWS);
      for (uint32_t r = r0; r < r1; ++r) {
        for (uint32_t x = 0; x < W; ++x) {
          double* px = accum + 4ull * ((size_t)r * W + x);
          double* sq = sum_sq ? sum_sq + 3ull * ((size_t)r * W + x) : nullptr;
          for (uint32_t si = 0; si < n; ++si) {
            f3 c;
            if (integ.sample(x, y0 + r, first + si, c)) {
              px[0] += (double)c[0];
              px[1] += (double)c[1];
              px[2] += (double)c[2];
              px[3] += 1.0;
              if (sq) {
                sq[0] += (double)c[0] * (double)c[0];
                sq[1] += (double)c[1] * (double)c[1];
                sq[2] += (double)c[2] * (double)c[2];
              }
            }
          }
        }
      }
    }
  };
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = (int)std::min<uint32_t>((uint32_t)nt, H / BAND_ROWS + 1);
  nt = std::max(nt, 1);
  if (nt <= 1) {
    work();
  } else {
    std::vector<std::thread> ts;
    for (int i = 0; i < nt; ++i) ts.emplace_back(work);
    for (auto& t : ts) t.join();
  }
  return 0;
}

// Samples [first, first+n) of the pixels in columns [x0, x1) and rows
// [y0, y1) only, into accum ((y1-y0) x (x1-x0) x 4) and sum_sq: per pixel
// the same f64 sums in sample order as oxrcpu_render; rows go to threads.
int oxrcpu_render_rect(oxrcpu_scene* h, const oxr_camera* cam, const oxr_pt_options* opts, uint32_t first, uint32_t n,
                       uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1, int threads, double* accum, double* sum_sq) {
  if (!h || !cam || !opts || !accum) return -1;
  uint32_t W = std::max<uint32_t>(opts->width, 1), H = std::max<uint32_t>(opts->height, 1);
  if (y1 > H || y0 >= y1 || x1 > W || x0 >= x1) return -1;
  Integrator integ{h->s, *cam, W, H, std::isfinite(opts->ambient) ? fmaxr(opts->ambient, 0.0f) : 0.0f, *opts};
  const uint32_t rows = y1 - y0, cols = x1 - x0;
  std::atomic<uint32_t> next{0};
  auto work = [&]() {
    for (;;) {
      uint32_t r = next.fetch_add(1);
      if (r >= rows) break;
      for (uint32_t c = 0; c < cols; ++c) {
        double* px = accum + 4ull * ((size_t)r * cols + c);
        double* sq = sum_sq ? sum_sq + 3ull * ((size_t)r * cols + c) : nullptr;
        for (uint32_t si = 0; si < n; ++si) {
          f3 v;
          if (integ.sample(x0 + c, y0 + r, first + si, v)) {
            for (int k = 0; k < 3; ++k) px[k] += (double)v[k];
            px[3] += 1.0;
            if (sq)
              for (int k = 0; k < 3; ++k) sq[k] += (double)v[k] * (double)v[k];
          }
        }
      }
    }
  };
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = std::max(1, (int)std::min<uint32_t>((uint32_t)nt, rows));
  if (nt <= 1) {
    work();
  } else {
    std::vector<std::thread> ts;
    for (int i = 0; i < nt; ++i) ts.emplace_back(work);
    for (auto& t : ts) t.join();
  }
  return 0;
}

// One sample of one pixel: writes rgb, returns 1 covered / 0 uncovered.
int oxrcpu_sample_one(oxrcpu_scene* h, const oxr_camera* cam, const oxr_pt_options* opts, uint32_t x, uint32_t y,
                      uint32_t index, float* rgb) {
  uint32_t W = std::max<uint32_t>(opts->width, 1), H = std::max<uint32_t>(opts->height, 1);
  Integrator integ{h->s, *cam, W, H, std::isfinite(opts->ambient) ? fmaxr(opts->ambient, 0.0f) : 0.0f, *opts};
  f3 c;
  bool cov = integ.sample(x, y, index, c);
  rgb[0] = cov ? c[0] : 0.0f;
  rgb[1] = cov ? c[1] : 0.0f;
  rgb[2] = cov ? c[2] : 0.0f;
  return cov ? 1 : 0;
}

// Closest hit of n rays. mode 0: BVH, 1: brute force. flags bit 0:
// primary back-face culling of single-sided materials.
int oxrcpu_trace_closest(oxrcpu_scene* h, uint32_t n, const oxr_ray* rays, uint32_t flags, int mode, oxr_hit* hits) {
  const Scene& s = h->s;
  bool cull = (flags & 1u) != 0;
  for (uint32_t i = 0; i < n; ++i) {
    const oxr_ray& r = rays[i];
    Hit hit;
    auto acc = [&](const Hit& c) { return !(cull && !c.front && !s.mat(c.g).double_sided); };
    bool got = closest_hit(s, {r.o[0], r.o[1], r.o[2]}, {r.d[0], r.d[1], r.d[2]}, r.t_min, r.t_max, acc, hit,
                           mode == 1);
    if (got) {
      hits[i].t = hit.t;
      hits[i].prim = hit.g | (hit.front ? 0u : OXR_HIT_BACK);
      hits[i].u = hit.bary[1];
      hits[i].v = hit.bary[2];
    } else {
      hits[i].t = 0.0f;
      hits[i].prim = OXR_HIT_MISS;
      hits[i].u = 0.0f;
      hits[i].v = 0.0f;
    }
  }
  return 0;
}

int oxrcpu_trace_any(oxrcpu_scene* h, uint32_t n, const oxr_ray* rays, int mode, uint8_t* occ) {
  const Scene& s = h->s;
  for (uint32_t i = 0; i < n; ++i) {
    const oxr_ray& r = rays[i];
    occ[i] = occluded(s, {r.o[0], r.o[1], r.o[2]}, {r.d[0], r.d[1], r.d[2]}, r.t_min, r.t_max,
                      [](const Hit&) { return true; }, mode == 1)
                 ? 1
                 : 0;
  }
  return 0;
}

void oxrcpu_sobol_owen_4d(uint32_t n, const uint32_t* idx, const uint32_t* seed, float* out) {
  for (uint32_t i = 0; i < n; ++i) {
    auto p = sobol_owen_4d(idx[i], seed[i]);
    for (int d = 0; d < 4; ++d) out[4 * i + d] = p[d];
  }
}

uint32_t oxrcpu_sobol(uint32_t index, int dim) { return sobol(index, dim); }
uint32_t oxrcpu_pcg_hash(uint32_t x) { return pcg_hash(x); }
void oxrcpu_sobol_matrices(uint32_t* out128) {
  for (int d = 0; d < 4; ++d)
    for (int k = 0; k < 32; ++k) out128[d * 32 + k] = sobol_matrices().m[d][k];
}
void oxrcpu_sheen_lut(float* out512) { std::memcpy(out512, sheen_lut().v, sizeof(float) * 512); }

static MaterialSample ms_from(const oxr_material_sample& m) {
  MaterialSample s;
  s.base_color[0] = m.base_color[0];
  s.base_color[1] = m.base_color[1];
  s.base_color[2] = m.base_color[2];
  s.base_color[3] = 1.0f;
  s.metallic = m.metallic;
  s.roughness = m.roughness;
  s.normal = {m.normal[0], m.normal[1], m.normal[2]};
  s.ior = m.ior;
  s.specular = m.specular;
  s.specular_color = {m.specular_color[0], m.specular_color[1], m.specular_color[2]};
  s.transmission = m.transmission;
  s.thickness = m.thickness;
  s.clearcoat = m.clearcoat;
  s.clearcoat_roughness = m.clearcoat_roughness;
  s.clearcoat_normal = {m.clearcoat_normal[0], m.clearcoat_normal[1], m.clearcoat_normal[2]};
  s.sheen_color = {m.sheen_color[0], m.sheen_color[1], m.sheen_color[2]};
  s.sheen_roughness = m.sheen_roughness;
  s.anisotropy = m.anisotropy;
  s.anisotropy_direction = {m.anisotropy_direction[0], m.anisotropy_direction[1], m.anisotropy_direction[2]};
  s.iridescence = m.iridescence;
  s.iridescence_ior = m.iridescence_ior;
  s.iridescence_thickness = m.iridescence_thickness;
  s.diffuse_transmission = m.diffuse_transmission;
  s.diffuse_transmission_color = {m.diffuse_transmission_color[0], m.diffuse_transmission_color[1],
                                  m.diffuse_transmission_color[2]};
  return s;
}

void oxrcpu_bsdf_eval(uint32_t n, const oxr_material_sample* ms, const float* v, const float* ng,
                      const uint8_t* front, const float* l, float* f, float* pdf) {
  for (uint32_t i = 0; i < n; ++i) {
    Bsdf b(ms_from(ms[i]), {v[3 * i], v[3 * i + 1], v[3 * i + 2]}, {ng[3 * i], ng[3 * i + 1], ng[3 * i + 2]},
           front[i] != 0);
    f3 ff;
    float p;
    b.eval({l[3 * i], l[3 * i + 1], l[3 * i + 2]}, ff, p);
    f[3 * i] = ff[0];
    f[3 * i + 1] = ff[1];
    f[3 * i + 2] = ff[2];
    pdf[i] = p;
  }
}

void oxrcpu_bsdf_sample(uint32_t n, const oxr_material_sample* ms, const float* v, const float* ng,
                        const uint8_t* front, const float* u, float* l, float* f, float* pdf, uint8_t* ok) {
  for (uint32_t i = 0; i < n; ++i) {
    Bsdf b(ms_from(ms[i]), {v[3 * i], v[3 * i + 1], v[3 * i + 2]}, {ng[3 * i], ng[3 * i + 1], ng[3 * i + 2]},
           front[i] != 0);
    f3 ll = {0, 0, 0}, ff = {0, 0, 0};
    float p = 0.0f;
    bool got = b.sample(&u[3 * i], ll, ff, p);
    for (int k = 0; k < 3; ++k) {
      l[3 * i + k] = got ? ll[k] : 0.0f;
      f[3 * i + k] = got ? ff[k] : 0.0f;
    }
    pdf[i] = got ? p : 0.0f;
    ok[i] = got ? 1 : 0;
  }
}

void oxrcpu_bsdf_probs(const oxr_material_sample* ms, const float* v, const float* ng, uint8_t front, float* out6) {
  Bsdf b(ms_from(*ms), {v[0], v[1], v[2]}, {ng[0], ng[1], ng[2]}, front != 0);
  for (int i = 0; i < LOBES; ++i) out6[i] = b.probs[i];
}

float oxrcpu_triangle_solid_angle(const float* p, const float* a, const float* b, const float* c) {
  return triangle_solid_angle({p[0], p[1], p[2]}, {a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]});
}

int oxrcpu_sample_spherical_triangle(const float* p, const float* a, const float* b, const float* c, float u0,
                                     float u1, float* out) {
  f3 o;
  bool ok = sample_spherical_triangle({p[0], p[1], p[2]}, {a[0], a[1], a[2]}, {b[0], b[1], b[2]},
                                      {c[0], c[1], c[2]}, u0, u1, o);
  if (ok) {
    out[0] = o[0];
    out[1] = o[1];
    out[2] = o[2];
  }
  return ok ? 1 : 0;
}

void oxrcpu_iridescent_fresnel(float outside, float film, const float* f0, float thick, float cos1, float* out) {
  f3 r = iridescent_fresnel(outside, film, {f0[0], f0[1], f0[2]}, thick, cos1);
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[2];
}

void oxrcpu_offset_ray_origin(const float* p, const float* n, float* out) {
  f3 r = offset_ray_origin({p[0], p[1], p[2]}, {n[0], n[1], n[2]});
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[2];
}

int oxrcpu_barycentric_of(const float* q, const float* p0, const float* p1, const float* p2, float* out) {
  f3 b;
  bool ok = barycentric_of({q[0], q[1], q[2]}, {p0[0], p0[1], p0[2]}, {p1[0], p1[1], p1[2]}, {p2[0], p2[1], p2[2]}, b);
  if (ok) {
    out[0] = b[0];
    out[1] = b[1];
    out[2] = b[2];
  }
  return ok ? 1 : 0;
}

// Surface at a hit: position, geometric normal, shading normal.
void oxrcpu_surface(oxrcpu_scene* h, uint32_t g, const float* bary, float* out9) {
  Surface sf = surface(h->s, g, {bary[0], bary[1], bary[2]});
  for (int k = 0; k < 3; ++k) {
    out9[k] = sf.position[k];
    out9[3 + k] = sf.geometric_normal[k];
    out9[6 + k] = sf.shading_normal[k];
  }
}

// PreparedTexture::sample_lod of texture `tex` in sRGB (srgb != 0) or
// linear space at n (u, v) pairs and levels of detail: 4 floats each.
int oxrcpu_tex_sample(oxrcpu_scene* h, uint32_t tex, int srgb, uint32_t n, const float* uv, const float* lod,
                      float* out) {
  if (tex >= h->s.texs.size()) return -1;
  for (uint32_t i = 0; i < n; ++i) {
    f4 r = sample_lod(h->s.texs[tex], srgb != 0, uv[2 * i], uv[2 * i + 1], lod[i]);
    for (int c = 0; c < 4; ++c) out[4 * i + c] = r[c];
  }
  return 0;
}

// PreparedTexture::sample_grad's level of detail for (du/dx, dv/dx,
// du/dy, dv/dy) quadruples.
int oxrcpu_tex_grad_lod(oxrcpu_scene* h, uint32_t tex, uint32_t n, const float* duv, float* lod) {
  if (tex >= h->s.texs.size()) return -1;
  for (uint32_t i = 0; i < n; ++i)
    lod[i] = grad_lod(h->s.texs[tex], duv[4 * i], duv[4 * i + 1], duv[4 * i + 2], duv[4 * i + 3]);
  return 0;
}

// Mip level `level` of a texture's chain: returns the level count; with
// `texels` non-null copies w*h*4 floats (w, h into wh[0..1]).
int oxrcpu_tex_level(oxrcpu_scene* h, uint32_t tex, int srgb, uint32_t level, uint32_t* wh, float* texels) {
  if (tex >= h->s.texs.size()) return -1;
  const std::vector<MipLevel>& c = h->s.texs[tex].chain[srgb ? 0 : 1];
  if (level < c.size()) {
    wh[0] = c[level].w;
    wh[1] = c[level].h;
    if (texels)
      for (size_t i = 0; i < c[level].t.size(); ++i)
        for (int k = 0; k < 4; ++k) texels[4 * i + k] = c[level].t[i][k];
  }
  return (int)c.size();
}

// TraceScene::material_oriented at n hits (triangle g, barycentrics b),
// with the ray cone (width, dir) when width > 0 and level 0 otherwise,
// and the double-sided flip (normals reversed back, as the integrator
// uses them); then the material's emission at level 0
// and the alpha of the any-hit filter. 43 floats per hit: the 35 of
// oxr_material_sample, base alpha, emissive (3), emission (3), alpha.
int oxrcpu_material_eval(oxrcpu_scene* h, uint32_t n, const uint32_t* g, const float* bary, const float* width,
                         const float* dir, const uint8_t* flip, float* out) {
  const Scene& s = h->s;
  oxr_camera cam{};
  oxr_pt_options opt{};
  Integrator it{s, cam, 1, 1, 0.0f, opt};
  for (uint32_t i = 0; i < n; ++i) {
    if (g[i] >= s.n) return -1;
    f3 b = {bary[3 * i], bary[3 * i + 1], bary[3 * i + 2]};
    TexLod L;
    if (width[i] > 0.0f) {
      L.cone = true;
      L.width = width[i];
      L.dir = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
    }
    Surface sf = surface(s, g[i], b);
    MaterialSample m = material_at(s, g[i], b, sf, flip[i] != 0, L);
    if (flip[i]) {  // the integrator's flip back (src/pathtrace.rs:2097-2101)
      m.normal = neg(m.normal);
      m.clearcoat_normal = neg(m.clearcoat_normal);
    }
    float* o = out + 43ull * i;
    float v[43] = {m.base_color[0], m.base_color[1], m.base_color[2], m.metallic, m.roughness,
                   m.normal[0], m.normal[1], m.normal[2], m.ior, m.specular,
                   m.specular_color[0], m.specular_color[1], m.specular_color[2], m.transmission, m.thickness,
                   m.clearcoat, m.clearcoat_roughness, m.clearcoat_normal[0], m.clearcoat_normal[1],
                   m.clearcoat_normal[2], m.sheen_color[0], m.sheen_color[1], m.sheen_color[2], m.sheen_roughness,
                   m.anisotropy, m.anisotropy_direction[0], m.anisotropy_direction[1], m.anisotropy_direction[2],
                   m.iridescence, m.iridescence_ior, m.iridescence_thickness, m.diffuse_transmission,
                   m.diffuse_transmission_color[0], m.diffuse_transmission_color[1],
                   m.diffuse_transmission_color[2], m.base_color[3], m.emissive[0], m.emissive[1], m.emissive[2],
                   0.0f, 0.0f, 0.0f, 0.0f};
    f3 e = it.emission(g[i], b);
    v[39] = e[0];
    v[40] = e[1];
    v[41] = e[2];
    v[42] = tri_alpha(s, g[i], b);
    for (int k = 0; k < 43; ++k) o[k] = v[k];
  }
  return 0;
}

// Resolve (PathTracer::hdr / image): accum W*H*4 doubles.
void oxrcpu_resolve_hdr(const double* accum, uint32_t npx, uint32_t samples, const oxr_display* disp, float* out) {
  float bg[4];
  for (int k = 0; k < 3; ++k) bg[k] = srgb_to_linear((float)disp->background[k] / 255.0f);
  bg[3] = (float)disp->background[3] / 255.0f;
  for (uint32_t i = 0; i < npx; ++i) {
    const double* a = accum + 4ull * i;
    float* o = out + 4ull * i;
    if (a[3] <= 0.0 || samples == 0) {
      for (int k = 0; k < 4; ++k) o[k] = bg[k];
      continue;
    }
    f3 m = {(float)(a[0] / a[3]), (float)(a[1] / a[3]), (float)(a[2] / a[3])};
    float f = (float)clampd(a[3] / (double)samples, 0.0, 1.0);
    if (f >= 1.0f) {
      o[0] = m[0];
      o[1] = m[1];
      o[2] = m[2];
      o[3] = 1.0f;
    } else {
      float a_sum = f + (1.0f - f) * bg[3];
      if (a_sum > 0.0f) {
        for (int k = 0; k < 3; ++k) o[k] = (f * m[k] + (1.0f - f) * bg[k] * bg[3]) / a_sum;
        o[3] = a_sum;
      } else {
        for (int k = 0; k < 3; ++k) o[k] = f * m[k] + (1.0f - f) * bg[k];
        o[3] = 0.0f;
      }
    }
  }
}

void oxrcpu_resolve_rgba8(const double* accum, uint32_t npx, uint32_t samples, const oxr_display* disp,
                          uint8_t* out) {
  float bg[4];
  for (int k = 0; k < 3; ++k) bg[k] = srgb_to_linear((float)disp->background[k] / 255.0f);
  bg[3] = (float)disp->background[3] / 255.0f;
  for (uint32_t i = 0; i < npx; ++i) {
    const double* a = accum + 4ull * i;
    uint8_t* o = out + 4ull * i;
    if (a[3] <= 0.0 || samples == 0) {
      for (int k = 0; k < 4; ++k) o[k] = disp->background[k];
      continue;
    }
    f3 m = {(float)(a[0] / a[3]), (float)(a[1] / a[3]), (float)(a[2] / a[3])};
    float f = (float)clampd(a[3] / (double)samples, 0.0, 1.0);
    f3 c = tone_map(disp->tone_map, scale(m, disp->exposure));
    float r[4];
    if (f >= 1.0f) {
      r[0] = c[0];
      r[1] = c[1];
      r[2] = c[2];
      r[3] = 1.0f;
    } else {
      float a_sum = f + (1.0f - f) * bg[3];
      if (a_sum > 0.0f) {
        for (int k = 0; k < 3; ++k) r[k] = (f * c[k] + (1.0f - f) * bg[k] * bg[3]) / a_sum;
        r[3] = a_sum;
      } else {
        for (int k = 0; k < 3; ++k) r[k] = f * c[k] + (1.0f - f) * bg[k];
        r[3] = 0.0f;
      }
    }
    for (int k = 0; k < 3; ++k) o[k] = linear_to_srgb_byte(r[k]);
    o[3] = (uint8_t)std::round(clampr(r[3], 0.0f, 1.0f) * 255.0f);
  }
}


// The display-transform pieces of the resolve, one value at a time, for the
// pins of src/hdr.rs:222-279: ToneMap::apply (src/hdr.rs:44-66),
// linear_to_srgb_byte (:116-129) and srgb_to_linear (:76-82).
void oxrcpu_tone_map(uint32_t tm, uint32_t n, const float* rgb, float* out) {
  for (uint32_t i = 0; i < n; ++i) {
    f3 c = tone_map(tm, {rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]});
    for (int k = 0; k < 3; ++k) out[3 * i + k] = c[k];
  }
}
void oxrcpu_linear_to_srgb_byte(uint32_t n, const float* c, uint8_t* out) {
  for (uint32_t i = 0; i < n; ++i) out[i] = linear_to_srgb_byte(c[i]);
}
void oxrcpu_srgb_to_linear(uint32_t n, const float* c, float* out) {
  for (uint32_t i = 0; i < n; ++i) out[i] = srgb_to_linear(c[i]);
}

// Records the closest-hit rays of samples [first, first+n) of rows
// [y0, y1) (single thread, row-major, sample-minor) into out (8 floats per
// ray: o, k, d, t_max), at most max_rays. Returns the number recorded.
static uint32_t record_rays(oxrcpu_scene* h, const oxr_camera* cam, const oxr_pt_options* opts, uint32_t first,
                            uint32_t n, uint32_t y0, uint32_t y1, uint32_t x_step, uint32_t max_rays, float* out,
                            bool shadow) {
  uint32_t W = std::max<uint32_t>(opts->width, 1), H = std::max<uint32_t>(opts->height, 1);
  std::vector<std::array<float, 8>> rec;
  Integrator integ{h->s, *cam, W, H, std::isfinite(opts->ambient) ? fmaxr(opts->ambient, 0.0f) : 0.0f, *opts};
  if (shadow)
    integ.record_shadow = &rec;
  else
    integ.record = &rec;
  for (uint32_t y = y0; y < std::min(y1, H); ++y)
    for (uint32_t x = 0; x < W; x += std::max(x_step, 1u))
      for (uint32_t si = 0; si < n; ++si) {
        f3 c;
        integ.sample(x, y, first + si, c);
        if (rec.size() >= max_rays) goto done;
      }
done:
  uint32_t m = (uint32_t)std::min<size_t>(rec.size(), max_rays);
  for (uint32_t i = 0; i < m; ++i)
    for (int k = 0; k < 8; ++k) out[8 * i + k] = rec[i][k];
  return m;
}
uint32_t oxrcpu_record_rays(oxrcpu_scene* h, const oxr_camera* cam, const oxr_pt_options* opts, uint32_t first,
                            uint32_t n, uint32_t y0, uint32_t y1, uint32_t x_step, uint32_t max_rays, float* out) {
  return record_rays(h, cam, opts, first, n, y0, y1, x_step, max_rays, out, false);
}
// The occlusion rays of the same paths (8 floats: o, skip id bits, d, t_max).
uint32_t oxrcpu_record_shadow_rays(oxrcpu_scene* h, const oxr_camera* cam, const oxr_pt_options* opts, uint32_t first,
                                   uint32_t n, uint32_t y0, uint32_t y1, uint32_t x_step, uint32_t max_rays,
                                   float* out) {
  return record_rays(h, cam, opts, first, n, y0, y1, x_step, max_rays, out, true);
}

// Traversal cost of rays on the device's flattened 4-wide BVH (64-byte
// nodes with 8-bit quantised child boxes, decoded as the device decodes
// them; 48-byte triangles: v0, gid, v1, flags, v2, alpha), restating the
// device's closest-hit order: leaves of a node first, inner children near
// to far.
// counts[2*i] = nodes fetched, counts[2*i+1] = triangles fetched;
// depth[i] (when not null) = the most entries the ray's traversal stack
// held at once (the device keeps 16 in LDS and spills the rest; the
// nearest child to visit next is held in a register, not on the stack).
// any: the occlusion order instead: the ray (o, skip id bits, d, t_max)
// ends at the first triangle hit in range (the skip id excluded), and the
// triangles counted are those tested up to it.
// Direction components below 2^-64 in magnitude (zero included) take the
// reciprocal +-2^64, as on the device: an infinite one turns the node's
// slab offset into NaN and the axis then culls nothing.
static float slab_rcp(float d) { return std::fabs(d) < 0x1p-64f ? std::copysign(0x1p64f, d) : 1.0f / d; }
static void bvh4_walk(const float* nodes, const float* tris, uint32_t n_rays, const float* rays, uint32_t* counts,
                      uint32_t* depth, bool any = false) {
  for (uint32_t r = 0; r < n_rays; ++r) {
    const float* ry = rays + 8ull * r;
    f3 o = {ry[0], ry[1], ry[2]}, d = {ry[4], ry[5], ry[6]};
    float tmin = 0.0f, tmax = ry[7];
    uint32_t skip;
    std::memcpy(&skip, &ry[3], 4);
    bool done_any = false;
    PRay pr = prepare_ray(o, d);
    f3 inv = {slab_rcp(d[0]), slab_rcp(d[1]), slab_rcp(d[2])};
    f3 oi = {o[0] * inv[0], o[1] * inv[1], o[2] * inv[2]};
    std::vector<uint32_t> stack;
    uint32_t node = 0, nn = 0, nt = 0, dmax = 0;
    uint32_t best = 0xffffffffu;
    for (;;) {
      ++nn;
      // 64-byte node: org.xyz, exponents; byte planes lo.x hi.x lo.y hi.y
      // lo.z hi.z and counts (child c in byte c); child refs at words 12-15.
      const float* N = nodes + 16ull * node;
      const uint32_t* NU = reinterpret_cast<const uint32_t*>(N);
      float tn[4];
      float sk[3], ok[3];
      for (int k = 0; k < 3; ++k) {
        sk[k] = std::ldexp(inv[k], (int)(int8_t)((NU[3] >> (8 * k)) & 0xffu));
        ok[k] = std::fma(N[k], inv[k], -oi[k]);
      }
      for (int c = 0; c < 4; ++c) {
        float t0[3], t1[3];
        for (int k = 0; k < 3; ++k) {
          const uint32_t wl = k < 2 ? NU[4 + 2 * k] : NU[8], wh = k < 2 ? NU[5 + 2 * k] : NU[9];
          t0[k] = std::fma((float)((wl >> (8 * c)) & 0xffu), sk[k], ok[k]);
          t1[k] = std::fma((float)((wh >> (8 * c)) & 0xffu), sk[k], ok[k]);
        }
        float tnear = tmin, tfar = tmax;
        for (int k = 0; k < 3; ++k) {
          bool neg = std::signbit(inv[k]);
          tnear = std::fmax(tnear, neg ? t1[k] : t0[k]);
          tfar = std::fmin(tfar, neg ? t0[k] : t1[k]);
        }
        bool hit = NU[12 + c] != 0xffffffffu && tnear <= tfar * (1.0f + 1.0f / 65536.0f);
        tn[c] = hit ? tnear : INF_F;
      }
      for (int c = 0; c < 4 && !done_any; ++c) {
        const uint32_t cnt = (NU[10] >> (8 * c)) & 0xffu;
        if (tn[c] != INF_F && cnt > 0) {
          for (uint32_t k = 0; k < cnt; ++k) {
            ++nt;
            const float* T = tris + 12ull * (NU[12 + c] + k);
            float v9[9] = {T[0], T[1], T[2], T[4], T[5], T[6], T[8], T[9], T[10]};
            float t, b1, b2;
            bool front;
            if (!intersect_tri(pr, v9, tmin, tmax, t, b1, b2, front)) continue;
            uint32_t gid = reinterpret_cast<const uint32_t*>(T)[3];
            if (any) {
              if (gid == skip) continue;
              done_any = true;
              break;
            }
            if (best != 0xffffffffu && !(t < tmax || (t == tmax && gid < best))) continue;
            tmax = t;
            best = gid;
          }
          tn[c] = INF_F;
        }
      }
      if (done_any) break;
      uint32_t ci[4] = {NU[12], NU[13], NU[14], NU[15]};
      for (int a = 0; a < 4; ++a)
        for (int b = a + 1; b < 4; ++b)
          if (tn[a] < tn[b]) {
            std::swap(tn[a], tn[b]);
            std::swap(ci[a], ci[b]);
          }
      size_t before = stack.size();
      for (int c = 0; c < 4; ++c)
        if (tn[c] != INF_F && tn[c] <= tmax * (1.0f + 1.0f / 65536.0f)) stack.push_back(ci[c]);
      // The device visits the nearest child next from a register and
      // stacks only the others.
      dmax = std::max(dmax, (uint32_t)(stack.size() - (stack.size() > before ? 1 : 0)));
      if (stack.empty()) break;
      node = stack.back();
      stack.pop_back();
    }
    counts[2 * r] = nn;
    counts[2 * r + 1] = nt;
    if (depth) depth[r] = dmax;
  }
}

void oxrcpu_bvh4_count(const float* nodes, const float* tris, uint32_t n_rays, const float* rays, uint32_t* counts) {
  bvh4_walk(nodes, tris, n_rays, rays, counts, nullptr);
}
void oxrcpu_bvh4_count_any(const float* nodes, const float* tris, uint32_t n_rays, const float* rays,
                           uint32_t* counts) {
  bvh4_walk(nodes, tris, n_rays, rays, counts, nullptr, true);
}
void oxrcpu_bvh4_stack_depth(const float* nodes, const float* tris, uint32_t n_rays, const float* rays,
                             uint32_t* depth) {
  std::vector<uint32_t> counts(2ull * std::max(n_rays, 1u));
  bvh4_walk(nodes, tris, n_rays, rays, counts.data(), depth);
}

}  // extern "C"
