// Device-side scalar/vector helpers, sampling primitives and the
// watertight triangle test of the oxr path tracer (gfx950).
//
// Float semantics follow oxideav-render's Rust code: no FMA contraction
// (the whole library builds with -ffp-contract=off), correctly rounded
// division and square root (HIP's default), `max`/`min` that ignore a NaN
// operand (f32::max / f32::min), `clamp` that keeps NaN.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define OXR_D __device__ __forceinline__

namespace oxr {

struct V3 {
  float x, y, z;
};

OXR_D V3 v3(float x, float y, float z) { return V3{x, y, z}; }
OXR_D V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
OXR_D V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
OXR_D V3 scale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
OXR_D V3 mul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
OXR_D V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
OXR_D float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
OXR_D V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
OXR_D float fmaxr(float a, float b) { return fmaxf(a, b); }
OXR_D float fminr(float a, float b) { return fminf(a, b); }
OXR_D float clampr(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
OXR_D double clampd(double x, double lo, double hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
// vec3_normalise: below f32::EPSILON the zero vector (math.rs:83-90).
OXR_D V3 normalise(V3 v) {
  float len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  if (len < 1.1920929e-7f) return v3(0.0f, 0.0f, 0.0f);
  return v3(v.x / len, v.y / len, v.z / len);
}
OXR_D float max3(V3 a) { return fmaxr(fmaxr(a.x, a.y), a.z); }
OXR_D float mean3(V3 a) { return (a.x + a.y + a.z) * (1.0f / 3.0f); }
OXR_D float luminance(V3 a) { return 0.2126f * a.x + 0.7152f * a.y + 0.0722f * a.z; }
OXR_D bool is_black(V3 a) { return !(a.x > 0.0f || a.y > 0.0f || a.z > 0.0f); }
OXR_D bool finite3(V3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
// powi(5) as LLVM expands it for Rust's f32::powi.
OXR_D float powi5(float x) {
  float x2 = x * x;
  float x4 = x2 * x2;
  return x * x4;
}
OXR_D float powi2(float x) { return x * x; }
OXR_D float comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
OXR_D V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }
OXR_D V3 f4xyz(float4 a) { return v3(a.x, a.y, a.z); }

constexpr float PI_F = 3.14159265358979323846f;
constexpr float FRAC_PI_2_F = 1.57079632679489661923f;
constexpr float FRAC_PI_4_F = 0.78539816339744830962f;

// ---------------------------------------------------------------------
// Sampling (pathtrace.rs:297-425): PCG hash, Owen-scrambled Sobol'.
// ---------------------------------------------------------------------
OXR_D uint32_t pcg_hash(uint32_t x) {
  uint32_t state = x * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
  return (word >> 22) ^ word;
}
OXR_D uint32_t hash_mix(uint32_t a, uint32_t b) { return pcg_hash(a ^ pcg_hash(b)); }

OXR_D uint32_t laine_karras(uint32_t x, uint32_t seed) {
  x += seed;
  x ^= x * 0x6c50b47cu;
  x ^= x * 0xb82f1e52u;
  x ^= x * 0xc7afe638u;
  x ^= x * 0x8d22f6e6u;
  return x;
}
OXR_D uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
  return __builtin_bitreverse32(laine_karras(__builtin_bitreverse32(x), seed));
}
OXR_D float to_unit(uint32_t v) { return (float)(v >> 8) * (1.0f / 16777216.0f); }

// Sobol' generator matrices of dimensions 1-3 (Joe-Kuo) as byte tables:
// c_sobol_tab[d-1][b][v] = XOR of the columns 8b..8b+7 selected by the
// bits of v. The matrix-vector product over GF(2) is then four
// independent lookups per dimension (12 KB, built by the host from the
// same recurrence); dimension 0 (van der Corput) is the bit reversal.
extern __constant__ uint32_t c_sobol_tab[3][4][256];

OXR_D uint32_t sobol(uint32_t index, int dim) {
  if (dim == 0) return __builtin_bitreverse32(index);
  const uint32_t(*t)[256] = c_sobol_tab[dim - 1];
  return t[0][index & 255u] ^ t[1][(index >> 8) & 255u] ^ t[2][(index >> 16) & 255u] ^ t[3][index >> 24];
}

struct F4 {
  float v[4];
};

OXR_D F4 sobol_owen_4d(uint32_t index, uint32_t pattern_seed) {
  uint32_t i = nested_uniform_scramble(index, pattern_seed);
  F4 out;
#pragma unroll
  for (int d = 0; d < 4; ++d)
    out.v[d] = to_unit(nested_uniform_scramble(sobol(i, d), hash_mix(pattern_seed, (uint32_t)d)));
  return out;
}

// Pixel/path seeds of Sampler::new (pathtrace.rs:409-416).
OXR_D uint32_t pixel_seed_of(uint32_t seed, uint32_t x, uint32_t y) { return hash_mix(hash_mix(seed, y), x); }
OXR_D F4 pattern(uint32_t pixel_seed, uint32_t index, uint32_t p) {
  return sobol_owen_4d(index, hash_mix(pixel_seed, p));
}
OXR_D float coin(uint32_t path_seed, uint32_t ray, uint32_t tri) {
  return to_unit(hash_mix(hash_mix(path_seed, ray), tri));
}

// ---------------------------------------------------------------------
// Trace-layer helpers (trace.rs:1120-1204).
// ---------------------------------------------------------------------
OXR_D bool barycentric_of(V3 q, V3 p0, V3 p1, V3 p2, V3& out) {
  V3 e1 = sub(p1, p0), e2 = sub(p2, p0), r = sub(q, p0);
  float d11 = dot(e1, e1), d12 = dot(e1, e2), d22 = dot(e2, e2);
  float r1 = dot(r, e1), r2 = dot(r, e2);
  float det = d11 * d22 - d12 * d12;
  if (!isfinite(det) || fabsf(det) <= 1.17549435e-38f) return false;
  float u = (d22 * r1 - d12 * r2) / det;
  float v = (d11 * r2 - d12 * r1) / det;
  out = v3(1.0f - u - v, u, v);
  return true;
}

OXR_D int32_t sat_i32(float x) {
  if (isnan(x)) return 0;
  if (x >= 2147483647.0f) return 2147483647;
  if (x <= -2147483648.0f) return (-2147483647 - 1);
  return (int32_t)x;
}

OXR_D float offset1(float p, float n) {
  int32_t of_i = sat_i32(256.0f * n);
  int32_t bits = __float_as_int(p);
  int32_t moved = p < 0.0f ? (int32_t)((uint32_t)bits - (uint32_t)of_i) : (int32_t)((uint32_t)bits + (uint32_t)of_i);
  float p_i = __int_as_float(moved);
  if (fabsf(p) < (1.0f / 32.0f)) return p + (1.0f / 65536.0f) * n;
  if (isfinite(p_i)) return p_i;
  return p;
}
// offset_ray_origin (trace.rs:1152-1175), Wächter & Binder 2019.
OXR_D V3 offset_ray_origin(V3 p, V3 n) { return v3(offset1(p.x, n.x), offset1(p.y, n.y), offset1(p.z, n.z)); }

// ---------------------------------------------------------------------
// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013).
// Barycentrics: b1 = V/det (weight of v1), b2 = W/det (weight of v2),
// v0's weight is 1 - b1 - b2. Front face <=> det > 0 (the ray runs
// against the CCW-front normal). Valid hits have t in [t_min, t_max].
// ---------------------------------------------------------------------
// The axis permutation (kx, ky, kz) is packed in one word, 2 bits each:
// the traversal loops keep it live across every triangle test, and three
// separate registers were spilled at 7 waves per SIMD.
struct WRay {
  V3 o;
  float Sx, Sy, Sz;
  uint32_t kp;  // kx | ky << 2 | kz << 4
  OXR_D int kx() const { return (int)(kp & 3u); }
  OXR_D int ky() const { return (int)((kp >> 2) & 3u); }
  OXR_D int kz() const { return (int)(kp >> 4); }
};

OXR_D WRay prepare_wray(V3 o, V3 d) {
  WRay r;
  r.o = o;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = 0;
  float m = ax;
  if (ay > m) {
    kz = 1;
    m = ay;
  }
  if (az > m) kz = 2;
  int kx = kz + 1 == 3 ? 0 : kz + 1;
  int ky = kx + 1 == 3 ? 0 : kx + 1;
  float dz = comp(d, kz);
  if (dz < 0.0f) {
    int t = kx;
    kx = ky;
    ky = t;
  }
  r.kp = (uint32_t)kx | ((uint32_t)ky << 2) | ((uint32_t)kz << 4);
  r.Sx = comp(d, kx) / dz;
  r.Sy = comp(d, ky) / dz;
  r.Sz = 1.0f / dz;
  return r;
}

OXR_D bool intersect_tri(const WRay& r, V3 v0, V3 v1, V3 v2, float t_min, float t_max, float& t_out, float& b1,
                         float& b2, bool& front) {
  V3 A = sub(v0, r.o), B = sub(v1, r.o), C = sub(v2, r.o);
  const int kx = r.kx(), ky = r.ky(), kz = r.kz();
  float Akz = comp(A, kz), Bkz = comp(B, kz), Ckz = comp(C, kz);
  float Ax = comp(A, kx) - r.Sx * Akz;
  float Ay = comp(A, ky) - r.Sy * Akz;
  float Bx = comp(B, kx) - r.Sx * Bkz;
  float By = comp(B, ky) - r.Sy * Bkz;
  float Cx = comp(C, kx) - r.Sx * Ckz;
  float Cy = comp(C, ky) - r.Sy * Ckz;
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
  float Az = r.Sz * Akz;
  float Bz = r.Sz * Bkz;
  float Cz = r.Sz * Ckz;
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

}  // namespace oxr
