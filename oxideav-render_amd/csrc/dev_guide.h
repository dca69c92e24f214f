// Device side of the learned path guide (gfx950): network layout, MFMA
// fragment conventions, hash-grid encoding and the vMF mixture used by the
// shading kernel. The design is stated in include/oxr/guide.h; the CPU
// oracle (oracle/guide_cpu.cpp) states the same arithmetic.
//
// MFMA conventions (v_mfma_f32_16x16x32_bf16, one wave = 16 queries):
//   A (16 x 32): lane l holds row l % 16, K slots 8 * (l / 16) + 0..7
//   B (32 x 16): lane l holds column l % 16, the same K slots
//   D (16 x 16): lane l holds column l % 16, rows 4 * (l / 16) + 0..3
// Activations live transposed: a layer computes Y^T = W X^T with the
// queries as the 16 columns, so lane (q = l % 16, g = l / 16) owns query q
// and, per 16-row output tile t, features 16t + 4g + 0..3. Those four
// values of tiles 2c and 2c+1 are exactly the K slots of lane group g in
// chunk c of the next layer, so activations pass from one layer to the
// next in registers with no shuffle; the weight fragments are packed in
// the matching K order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_math.h"

namespace oxr {

constexpr int G_LEVELS = 16;
constexpr int G_HID = 64;
constexpr int G_OUT = 48;
constexpr int G_LOBES = 8;
constexpr int G_NIN = 32;

// Offsets of the MLP part of the parameter array ([out][in] weights).
constexpr int G_OW1 = 0, G_OB1 = G_OW1 + G_HID * G_NIN, G_OW2 = G_OB1 + G_HID, G_OB2 = G_OW2 + G_HID * G_HID,
              G_OW3 = G_OB2 + G_HID, G_OB3 = G_OW3 + G_HID * G_HID, G_OW4 = G_OB3 + G_HID,
              G_OB4 = G_OW4 + G_OUT * G_HID, G_NMLP = G_OB4 + G_OUT;

// Fragment table: 64 lanes x 16 bytes (8 bf16) per fragment.
//   forward  L1: 4 (tile)          L2, L3: 8 (tile * 2 + chunk)   L4: 6
//   backward L4: 8 (row tile * 2 + chunk)  L3, L2: 8   L1: 4
constexpr int FW1 = 0, FW2 = 4, FW3 = 12, FW4 = 20, BW4 = 26, BW3 = 34, BW2 = 42, BW1 = 50, NFRAG = 54;
constexpr int NFRAG_FWD = 26;

struct DevGuideNet {
  const float* grid;   // [levels][T][2]
  const uint4* frag;   // NFRAG * 64 bf16 fragments
  const float* mlp;    // f32 master parameters (biases are read from here)
  uint32_t res[G_LEVELS];
  uint32_t dense_mask; // bit l: level l indexes densely
  uint32_t tmask;      // T - 1
  uint32_t log2_table;
  float lo[3], inv[3];
};

// Gradient copies of the coarse dense levels 0 .. nlev-1 (nlev <=
// G_COPY_LEVELS): G_GRAD_COPIES copies of `stride` floats; level l's
// dense table ((res_l + 1)^3 entries x 2 features) starts at float off[l]
// of each copy. Levels qualify while (res_l + 1)^3 <= G_COPY_MAX_ENTRIES.
constexpr int G_GRAD_COPIES = 8;
constexpr int G_COPY_LEVELS = 4;
constexpr uint32_t G_COPY_MAX_ENTRIES = 1u << 15;

struct GradCopies {
  float* p;
  uint32_t nlev;
  uint32_t stride;
  uint32_t off[G_COPY_LEVELS];
};

// Per-query guide output of the inference kernel: 8 x (mu, kappa), then
// the 8 lobe weights. 160 bytes.
struct GuideLobes {
  float4 mk[G_LOBES];
  float pi[G_LOBES];
};

// What the shading kernel needs of an attached guide for one bounce. The
// guide serves bounces k < kmax <= G_KREC (inference runs only for those).
// Training records: for paths with pid % 2^shift == 0, vertex k < G_KREC
// keeps 4 float4 at rec[((pid >> shift) * G_KREC + k) * 4]:
//   (position, mixture pdf), (direction, 1 = valid), (throughput after
//   the vertex, 0), (path radiance once the vertex's own NEE is in, 0).
constexpr int G_KREC = 4;  // = OXR_GUIDE_MAX_BOUNCES
constexpr uint32_t G_PATTERN = 0xA0000000u;  // Sobol' pattern id base of the guide draws

struct DevGuideUse {
  const GuideLobes* lobes;  // per entry of the bounce's ray queue
  float4* rec;              // null: no training records
  float mix;                // guide probability at eligible vertices
  uint32_t shift;
  uint32_t kmax;            // vertices k < kmax use the guide
};

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

OXR_D uint32_t bf16_bits(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
OXR_D float bf16_value(uint32_t b) { return __uint_as_float(b << 16); }
OXR_D uint4 pack8(const float v[8]) {
  return make_uint4(bf16_bits(v[0]) | (bf16_bits(v[1]) << 16), bf16_bits(v[2]) | (bf16_bits(v[3]) << 16),
                    bf16_bits(v[4]) | (bf16_bits(v[5]) << 16), bf16_bits(v[6]) | (bf16_bits(v[7]) << 16));
}
OXR_D uint4 pack44(f32x4_t a, f32x4_t b) {
  float v[8] = {a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
  return pack8(v);
}
OXR_D f32x4_t mfma16(uint4 a, uint4 b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c,
                                                 0, 0, 0);
}

// Output slot of (lobe, component) — the decoding rule of the header.
OXR_D constexpr int g_out_index(int lobe, int comp) {
  return 16 * ((5 * (lobe % 2) + comp) / 4) + 4 * (lobe / 2) + (5 * (lobe % 2) + comp) % 4;
}

// Corner indices / trilinear weights of one level of resolution R at
// position p (device order: corner bit 0 = x, bit 1 = y, bit 2 = z).
OXR_D void g_corners(const DevGuideNet& n, uint32_t R, bool dense, V3 p, uint32_t idx[8], float w[8]) {
  float Rf = (float)R;
  float pc[3] = {p.x, p.y, p.z};
  uint32_t base[3];
  float fr[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float x = (pc[d] - n.lo[d]) * n.inv[d];
    x = x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f;
    float s = x * Rf;
    int i = (int)floorf(s);
    if (i > (int)R - 1) i = (int)R - 1;
    base[d] = (uint32_t)i;
    fr[d] = s - (float)i;
  }
  uint32_t r = R + 1u;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    uint32_t cx = base[0] + (c & 1), cy = base[1] + ((c >> 1) & 1), cz = base[2] + ((c >> 2) & 1);
    idx[c] = dense ? cx + r * (cy + r * cz) : ((cx * 1u) ^ (cy * 2654435761u) ^ (cz * 805459861u)) & n.tmask;
    float wx = (c & 1) ? fr[0] : 1.0f - fr[0];
    float wy = (c & 2) ? fr[1] : 1.0f - fr[1];
    float wz = (c & 4) ? fr[2] : 1.0f - fr[2];
    w[c] = wx * wy * wz;
  }
}

// ---------------------------------------------------------------------
// vMF mixture: density and sampling (the shading kernel's proposal).
// ---------------------------------------------------------------------
OXR_D float vmf_pdf(float kappa, float c) {
  return kappa / (2.0f * PI_F * -expm1f(-2.0f * kappa)) * expf(kappa * (c - 1.0f));
}

OXR_D float guide_pdf(const GuideLobes* __restrict__ g, V3 w) {
  float q = 0.0f;
#pragma unroll 1
  for (int k = 0; k < G_LOBES; ++k) {
    float4 m = g->mk[k];
    q += g->pi[k] * vmf_pdf(m.w, m.x * w.x + m.y * w.y + m.z * w.z);
  }
  return q;
}

// Lobe by the running sum of pi, then w = 1 + log1p(-(1-u1)(1-e^{-2k}))/k
// and phi = 2 pi u2 in the Duff et al. frame of mu.
OXR_D V3 guide_sample(const GuideLobes* __restrict__ g, float u0, float u1, float u2) {
  int pick = G_LOBES - 1;
  bool found = false;
  float c = 0.0f;
#pragma unroll 1
  for (int k = 0; k < G_LOBES; ++k) {
    c += g->pi[k];
    if (!found && u0 < c) {
      pick = k;
      found = true;
    }
  }
  float4 m = g->mk[pick];
  float kap = m.w;
  float w = 1.0f + log1pf(-(1.0f - u1) * -expm1f(-2.0f * kap)) / kap;
  w = w < -1.0f ? -1.0f : (w > 1.0f ? 1.0f : w);
  float st = sqrtf(fmaxf(0.0f, 1.0f - w * w));
  float phi = 2.0f * PI_F * u2;
  float sign = copysignf(1.0f, m.z);
  float a = -1.0f / (sign + m.z);
  float bb = m.x * m.y * a;
  V3 t = v3(1.0f + sign * m.x * m.x * a, sign * bb, -sign * m.x);
  V3 b = v3(bb, sign + m.y * m.y * a, -m.y);
  float cx = st * cosf(phi), cy = st * sinf(phi);
  return v3(t.x * cx + b.x * cy + m.x * w, t.y * cx + b.y * cy + m.y * w, t.z * cx + b.z * cy + m.z * w);
}

}  // namespace oxr
