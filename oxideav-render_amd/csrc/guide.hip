// HIP kernels of the learned path guide (gfx950): hash-grid encoding +
// BF16 MFMA MLP inference, forward/backward over a training batch, the
// split-K weight-gradient GEMM, its fixed-order reduction, Adam, and the
// packing of bf16 weight fragments. Conventions are in dev_guide.h.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dev_guide.h"
#include "dev_scene.h"
#include "guide_rt.h"

namespace oxr {

constexpr int GUIDE_TB = 256;  // 4 waves; each wave owns 16 queries at a time
constexpr int N_BIAS = 3 * G_HID + G_OUT;

// Per-group offsets (bf16 elements) of the stored activations X0..X3 and
// deltas D1..D4, each [feature][16 queries].
constexpr int XS_GROUP = (G_NIN + 3 * G_HID) * 16;
constexpr int XO0 = 0, XO1 = G_NIN * 16, XO2 = XO1 + G_HID * 16, XO3 = XO2 + G_HID * 16;
constexpr int DS_GROUP = (3 * G_HID + G_OUT) * 16;
constexpr int DO1 = 0, DO2 = G_HID * 16, DO3 = DO2 + G_HID * 16, DO4 = DO3 + G_HID * 16;
constexpr int N_WTILES = 8 + 16 + 16 + 12;

template <int NF>
struct GuideShared {
  uint4 frag[NF * 64];
  float bias[N_BIAS];
  uint32_t res[G_LEVELS];
};

template <int NF>
OXR_D void stage(const DevGuideNet& n, GuideShared<NF>& sh) {
  for (int i = threadIdx.x; i < NF * 64; i += blockDim.x) sh.frag[i] = n.frag[i];
  for (int i = threadIdx.x; i < N_BIAS; i += blockDim.x) {
    int o = i < 3 * G_HID ? (i / G_HID) * (G_OB2 - G_OB1) + G_OB1 + i % G_HID : G_OB4 + (i - 3 * G_HID);
    sh.bias[i] = n.mlp[o];
  }
  if (threadIdx.x < G_LEVELS) sh.res[threadIdx.x] = n.res[threadIdx.x];
  __syncthreads();
}

OXR_D f32x4_t bias4(const float* b, int o0) { return f32x4_t{b[o0], b[o0 + 1], b[o0 + 2], b[o0 + 3]}; }
OXR_D f32x4_t relu4(f32x4_t a) {
  return f32x4_t{fmaxf(a[0], 0.0f), fmaxf(a[1], 0.0f), fmaxf(a[2], 0.0f), fmaxf(a[3], 0.0f)};
}
OXR_D uint32_t posmask4(f32x4_t a, int t) {
  return ((a[0] > 0.0f) | ((a[1] > 0.0f) << 1) | ((a[2] > 0.0f) << 2) | ((a[3] > 0.0f) << 3)) << (4 * t);
}
OXR_D f32x4_t masked(f32x4_t a, uint32_t m, int t) {
  m >>= 4 * t;
  return f32x4_t{(m & 1) ? a[0] : 0.0f, (m & 2) ? a[1] : 0.0f, (m & 4) ? a[2] : 0.0f, (m & 8) ? a[3] : 0.0f};
}
OXR_D f32x4_t round4(f32x4_t a) {
  return f32x4_t{bf16_value(bf16_bits(a[0])), bf16_value(bf16_bits(a[1])), bf16_value(bf16_bits(a[2])),
                 bf16_value(bf16_bits(a[3]))};
}

// Encoded features of levels 4g .. 4g+3 (this lane's K slots of layer 1).
OXR_D void encode4(const DevGuideNet& n, const uint32_t* res, V3 p, int g, float feat[8]) {
  const float2* grid2 = reinterpret_cast<const float2*>(n.grid);
  size_t T = (size_t)n.tmask + 1u;
#pragma unroll
  for (int lv = 0; lv < 4; ++lv) {
    int l = 4 * g + lv;
    uint32_t idx[8];
    float w[8];
    g_corners(n, res[l], (n.dense_mask >> l) & 1u, p, idx, w);
    float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      float2 e = grid2[(size_t)l * T + idx[c]];
      f0 += w[c] * e.x;
      f1 += w[c] * e.y;
    }
    feat[2 * lv] = f0;
    feat[2 * lv + 1] = f1;
  }
}

// Forward through the MLP. Returns the last layer's 3 output tiles;
// optionally keeps the bf16 activations and ReLU masks for backward.
struct Acts {
  uint4 x0, x1[2], x2[2], x3[2];
  uint32_t m1, m2, m3;
};

OXR_D void mlp_forward(const uint4* F, const float* bias, int lane, int g, uint4 x0, Acts& A, f32x4_t out[3]) {
  A.x0 = x0;
  f32x4_t a[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) a[t] = mfma16(F[(FW1 + t) * 64 + lane], x0, bias4(bias, 16 * t + 4 * g));
  A.m1 = posmask4(a[0], 0) | posmask4(a[1], 1) | posmask4(a[2], 2) | posmask4(a[3], 3);
  A.x1[0] = pack44(relu4(a[0]), relu4(a[1]));
  A.x1[1] = pack44(relu4(a[2]), relu4(a[3]));
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    a[t] = bias4(bias, G_HID + 16 * t + 4 * g);
    a[t] = mfma16(F[(FW2 + 2 * t) * 64 + lane], A.x1[0], a[t]);
    a[t] = mfma16(F[(FW2 + 2 * t + 1) * 64 + lane], A.x1[1], a[t]);
  }
  A.m2 = posmask4(a[0], 0) | posmask4(a[1], 1) | posmask4(a[2], 2) | posmask4(a[3], 3);
  A.x2[0] = pack44(relu4(a[0]), relu4(a[1]));
  A.x2[1] = pack44(relu4(a[2]), relu4(a[3]));
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    a[t] = bias4(bias, 2 * G_HID + 16 * t + 4 * g);
    a[t] = mfma16(F[(FW3 + 2 * t) * 64 + lane], A.x2[0], a[t]);
    a[t] = mfma16(F[(FW3 + 2 * t + 1) * 64 + lane], A.x2[1], a[t]);
  }
  A.m3 = posmask4(a[0], 0) | posmask4(a[1], 1) | posmask4(a[2], 2) | posmask4(a[3], 3);
  A.x3[0] = pack44(relu4(a[0]), relu4(a[1]));
  A.x3[1] = pack44(relu4(a[2]), relu4(a[3]));
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    out[t] = bias4(bias, 3 * G_HID + 16 * t + 4 * g);
    out[t] = mfma16(F[(FW4 + 2 * t) * 64 + lane], A.x3[0], out[t]);
    out[t] = mfma16(F[(FW4 + 2 * t + 1) * 64 + lane], A.x3[1], out[t]);
  }
}

// Lobe decoding of this lane's two lobes (2g, 2g+1) from its 12 output
// slots v[4t + j]; the softmax runs across the wave's four lane groups.
struct LanePair {
  V3 mu[2];
  float len[2], s[2], kappa[2], pi[2], logpi[2];
};

OXR_D void decode_pair(const float v[12], LanePair& P) {
  float mx = fmaxf(v[4], v[9]);
  mx = fmaxf(mx, __shfl_xor(mx, 16));
  mx = fmaxf(mx, __shfl_xor(mx, 32));
  float e0 = expf(v[4] - mx), e1 = expf(v[9] - mx);
  float sum = e0 + e1;
  sum += __shfl_xor(sum, 16);
  sum += __shfl_xor(sum, 32);
  float lsum = logf(sum);
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    int b = 5 * a;
    float m0 = v[b], m1 = v[b + 1], m2 = v[b + 2];
    float len = sqrtf(m0 * m0 + m1 * m1 + m2 * m2);
    P.len[a] = len;
    P.mu[a] = len > 1.0e-12f ? v3(m0 / len, m1 / len, m2 / len) : v3(0.0f, 0.0f, 1.0f);
    float s = v[b + 3];
    P.s[a] = s;
    float sc = s < -3.0f ? -3.0f : (s > 8.0f ? 8.0f : s);
    P.kappa[a] = expf(sc);
    P.pi[a] = (a == 0 ? e0 : e1) / sum;
    P.logpi[a] = v[b + 4] - mx - lsum;
  }
}

OXR_D V3 query_pos(int SRC, uint32_t i, const float* pos, const float4* qray, const uint4* hits) {
  if (SRC == 0) return ld3(pos + 3ull * i);
  float4 o = qray[2ull * i];
  float4 d = qray[2ull * i + 1];
  uint4 h = hits[i];
  float t = h.y == HIT_MISS ? 0.0f : __uint_as_float(h.x);
  return v3(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
}

// ---------------------------------------------------------------------
// Inference: positions (host array or the closest-hit queue: o + t d) ->
// per-query lobes (and optionally the 48 raw outputs).
// ---------------------------------------------------------------------
template <int SRC>
__global__ __launch_bounds__(GUIDE_TB) __attribute__((amdgpu_waves_per_eu(4))) void k_guide_infer(DevGuideNet n, uint32_t fixed_count,
                                                          const uint32_t* __restrict__ count_ptr,
                                                          const float* __restrict__ pos,
                                                          const float4* __restrict__ qray,
                                                          const uint4* __restrict__ hits, GuideLobes* __restrict__ out,
                                                          float* __restrict__ raw) {
  __shared__ GuideShared<NFRAG_FWD> sh;
  stage<NFRAG_FWD>(n, sh);
  const uint32_t count = count_ptr ? *count_ptr : fixed_count;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane & 15, g = lane >> 4;
  const uint32_t ngroups = (count + 15u) / 16u;
  for (uint32_t grp = blockIdx.x * (GUIDE_TB / 64) + wave; grp < ngroups; grp += gridDim.x * (GUIDE_TB / 64)) {
    uint32_t i = grp * 16u + q;
    bool valid = i < count;
    // Queue entries that missed need no lobes (shading never reads them):
    // a group whose 16 queries all missed is skipped.
    if (SRC == 1 && !__any(valid && hits[i].y != HIT_MISS)) continue;
    V3 p = valid ? query_pos(SRC, i, pos, qray, hits) : v3(0.0f, 0.0f, 0.0f);
    float feat[8];
    encode4(n, sh.res, p, g, feat);
    // An opaque copy of the lane id keeps the (loop-invariant) LDS
    // fragment loads inside the loop instead of pinning 26 fragments in
    // registers for the whole kernel.
    int ln = lane;
    asm volatile("" : "+v"(ln));
    Acts A;
    f32x4_t o4[3];
    mlp_forward(sh.frag, sh.bias, ln, g, pack8(feat), A, o4);
    float v[12];
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) v[4 * t + j] = o4[t][j];
    LanePair P;
    decode_pair(v, P);
    if (valid) {
      GuideLobes* L = out + i;
      L->mk[2 * g] = make_float4(P.mu[0].x, P.mu[0].y, P.mu[0].z, P.kappa[0]);
      L->mk[2 * g + 1] = make_float4(P.mu[1].x, P.mu[1].y, P.mu[1].z, P.kappa[1]);
      L->pi[2 * g] = P.pi[0];
      L->pi[2 * g + 1] = P.pi[1];
      if (raw) {
#pragma unroll
        for (int t = 0; t < 3; ++t)
#pragma unroll
          for (int j = 0; j < 4; ++j) raw[(size_t)i * G_OUT + 16 * t + 4 * g + j] = v[4 * t + j];
      }
    }
  }
}

// ---------------------------------------------------------------------
// Training: 1/sum(w), then forward + loss gradient + backward per group
// of 16 samples; activations/deltas (bf16) to xs/ds for the weight
// gradient; hash-grid gradients by f32 atomics.
// Sample i: samp[2i] = (pos, w), samp[2i+1] = (dir, proposal pdf; unused
// here). count % 32 == 0.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_guide_wsum(uint32_t count, const float4* __restrict__ samp,
                                                     float* __restrict__ inv_wsum, float* __restrict__ loss) {
  __shared__ float red[1024];
  float s = 0.0f;
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) s += samp[2ull * i].w;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 512; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *inv_wsum = red[0] > 0.0f ? 1.0f / red[0] : 0.0f;
    *loss = 0.0f;
  }
}

OXR_D void store4(uint16_t* base, int f0, int q, f32x4_t a) {
#pragma unroll
  for (int j = 0; j < 4; ++j) base[(f0 + j) * 16 + q] = (uint16_t)bf16_bits(a[j]);
}

// The first gc.nlev levels (coarse and dense: few entries, each hit by
// many samples of a batch) take their gradient atomics in G_GRAD_COPIES
// copies, chosen by block, which k_guide_fold then sums into ggrid:
// atomics to one address serialise, and on a single table these levels
// took most of the atomic time.
__global__ __launch_bounds__(GUIDE_TB) __attribute__((amdgpu_waves_per_eu(2))) void k_guide_fwdbwd(DevGuideNet n, uint32_t count,
                                                           const float4* __restrict__ samp,
                                                           const float* __restrict__ inv_wsum,
                                                           float* __restrict__ ggrid, GradCopies gc,
                                                           uint16_t* __restrict__ xs,
                                                           uint16_t* __restrict__ ds, float* __restrict__ loss) {
  __shared__ GuideShared<NFRAG> sh;
  stage<NFRAG>(n, sh);
  const float iw = *inv_wsum;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane & 15, g = lane >> 4;
  const uint32_t ngroups = count / 16u;
  const uint4* F = sh.frag;
  const size_t T = (size_t)n.tmask + 1u;
  float lacc = 0.0f;
  for (uint32_t grp = blockIdx.x * (GUIDE_TB / 64) + wave; grp < ngroups; grp += gridDim.x * (GUIDE_TB / 64)) {
    uint32_t i = grp * 16u + q;
    float4 pw = samp[2ull * i];
    float4 dd = samp[2ull * i + 1];
    V3 p = v3(pw.x, pw.y, pw.z), dir = v3(dd.x, dd.y, dd.z);
    float wn = pw.w * iw;
    float feat[8];
    encode4(n, sh.res, p, g, feat);
    uint4 x0 = pack8(feat);
    int ln = lane;
    asm volatile("" : "+v"(ln));
    Acts A;
    f32x4_t o4[3];
    mlp_forward(sh.frag, sh.bias, ln, g, x0, A, o4);
    float v[12];
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) v[4 * t + j] = o4[t][j];
    LanePair P;
    decode_pair(v, P);
    // Loss -wn * log q and d loss / d raw for this lane's lobes.
    float lp[2], c[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      c[a] = dot(P.mu[a], dir);
      float kap = P.kappa[a];
      lp[a] = P.logpi[a] + (logf(kap) - logf(2.0f * PI_F) - logf(-expm1f(-2.0f * kap)) + kap * (c[a] - 1.0f));
    }
    float mx = fmaxf(lp[0], lp[1]);
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    float se = expf(lp[0] - mx) + expf(lp[1] - mx);
    se += __shfl_xor(se, 16);
    se += __shfl_xor(se, 32);
    float lse = mx + logf(se);
    if (g == 0) lacc += -wn * lse;
    float dv[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) dv[k] = 0.0f;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      int b = 5 * a;
      float r = expf(lp[a] - lse);
      float kap = P.kappa[a];
      dv[b + 4] = -wn * (r - P.pi[a]);
      if (P.s[a] > -3.0f && P.s[a] < 8.0f)
        dv[b + 3] = -wn * r * (1.0f - 2.0f * kap / expm1f(2.0f * kap) + kap * (c[a] - 1.0f));
      if (P.len[a] > 1.0e-12f) {
        float k2 = -wn * r * kap;
        dv[b + 0] = k2 * (dir.x - c[a] * P.mu[a].x) / P.len[a];
        dv[b + 1] = k2 * (dir.y - c[a] * P.mu[a].y) / P.len[a];
        dv[b + 2] = k2 * (dir.z - c[a] * P.mu[a].z) / P.len[a];
      }
    }
    // Stored activations (bf16) for the weight gradient.
    uint16_t* X = xs + (size_t)grp * XS_GROUP;
    {
      const uint32_t* xw = reinterpret_cast<const uint32_t*>(&A.x0);
#pragma unroll
      for (int j = 0; j < 8; ++j) X[XO0 + (8 * g + j) * 16 + q] = (uint16_t)(xw[j >> 1] >> (16 * (j & 1)));
      const uint4* xl[3] = {A.x1, A.x2, A.x3};
      const int xo[3] = {XO1, XO2, XO3};
#pragma unroll
      for (int L = 0; L < 3; ++L)
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
          const uint32_t* w32 = reinterpret_cast<const uint32_t*>(&xl[L][cc]);
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            int f = 32 * cc + 16 * (j >> 2) + 4 * g + (j & 3);
            X[xo[L] + f * 16 + q] = (uint16_t)(w32[j >> 1] >> (16 * (j & 1)));
          }
        }
    }
    uint16_t* D = ds + (size_t)grp * DS_GROUP;
    // Layer 4 deltas.
    f32x4_t d4[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) d4[t] = round4(f32x4_t{dv[4 * t], dv[4 * t + 1], dv[4 * t + 2], dv[4 * t + 3]});
#pragma unroll
    for (int t = 0; t < 3; ++t) store4(D + DO4, 16 * t + 4 * g, q, d4[t]);
    uint4 db[2] = {pack44(d4[0], d4[1]), pack44(d4[2], f32x4_t{0.0f, 0.0f, 0.0f, 0.0f})};
    // Back through W4, W3, W2 (ReLU masks of the forward pass).
    const int bw[3] = {BW4, BW3, BW2};
    const int dofs[3] = {DO3, DO2, DO1};
    const uint32_t masks[3] = {A.m3, A.m2, A.m1};
#pragma unroll
    for (int L = 0; L < 3; ++L) {
      f32x4_t dx[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        dx[t] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
        dx[t] = mfma16(F[(bw[L] + 2 * t) * 64 + ln], db[0], dx[t]);
        dx[t] = mfma16(F[(bw[L] + 2 * t + 1) * 64 + ln], db[1], dx[t]);
        dx[t] = round4(masked(dx[t], masks[L], t));
        store4(D + dofs[L], 16 * t + 4 * g, q, dx[t]);
      }
      db[0] = pack44(dx[0], dx[1]);
      db[1] = pack44(dx[2], dx[3]);
    }
    // Back through W1 to the encoded features of levels 4g .. 4g+3.
    f32x4_t dx0[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      dx0[t] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
      dx0[t] = mfma16(F[(BW1 + 2 * t) * 64 + ln], db[0], dx0[t]);
      dx0[t] = mfma16(F[(BW1 + 2 * t + 1) * 64 + ln], db[1], dx0[t]);
    }
    // Grid gradient. Every lane of the wave runs this (the merge below reads
    // the other lanes of its group); a zero-weight sample's gradient is 0.
#pragma unroll
    for (int lv = 0; lv < 4; ++lv) {
      int l = 4 * g + lv;
      float df0 = dx0[lv >> 1][2 * (lv & 1)], df1 = dx0[lv >> 1][2 * (lv & 1) + 1];
      uint32_t idx[8];
      float w[8];
      g_corners(n, sh.res[l], (n.dense_mask >> l) & 1u, p, idx, w);
      // Copied levels are below G_COPY_LEVELS <= 4, so only lane group 0
      // (levels 0..3, l == lv) can meet one and the offset index is a
      // constant.
      float* gl = (g == 0 && (uint32_t)lv < gc.nlev)
                      ? gc.p + (size_t)(blockIdx.x % G_GRAD_COPIES) * gc.stride + gc.off[lv]
                      : ggrid + (size_t)l * T * 2;
#pragma unroll
      for (int cc = 0; cc < 8; ++cc) {
        // The group's 16 samples whose corner cc falls on the same entry of
        // level l: the lowest of those lanes adds their sum, the others
        // nothing (same-address atomics serialise; samples of one region
        // share the coarse and middle levels' entries).
        const float c0 = w[cc] * df0, c1 = w[cc] * df1;
        float s0 = c0, s1 = c1;
        bool lead = true;
#pragma unroll
        for (int sft = 1; sft < 16; ++sft) {
          const int oq = (q + sft) & 15;
          const int src = (lane & ~15) | oq;
          const uint32_t oi = (uint32_t)__shfl((int)idx[cc], src);
          const float o0 = __shfl(c0, src), o1 = __shfl(c1, src);
          if (oi == idx[cc]) {
            s0 += o0;
            s1 += o1;
            lead = lead && oq > q;
          }
        }
        float* e = gl + (size_t)idx[cc] * 2;
        if (lead && s0 != 0.0f) atomicAdd(e, s0);
        if (lead && s1 != 0.0f) atomicAdd(e + 1, s1);
      }
    }
  }
  // One atomic per wave for the loss statistic.
  for (int off = 32; off > 0; off >>= 1) lacc += __shfl_xor(lacc, off);
  if (lane == 0 && lacc != 0.0f) atomicAdd(loss, lacc);
}

// ---------------------------------------------------------------------
// Weight gradient dW[o][i] = sum_q dZ[o][q] X[i][q] as a split-K GEMM:
// block (tile, split) = one wave, K = its range of 32-sample pairs.
// Tiles: L1 4x2, L2 4x4, L3 4x4, L4 3x4 (output tile, input tile).
// ---------------------------------------------------------------------
OXR_D void wtile(int tile, int& layer, int& to, int& ti) {
  if (tile < 8) {
    layer = 0;
    to = tile >> 1;
    ti = tile & 1;
  } else if (tile < 24) {
    layer = 1;
    to = (tile - 8) >> 2;
    ti = (tile - 8) & 3;
  } else if (tile < 40) {
    layer = 2;
    to = (tile - 24) >> 2;
    ti = (tile - 24) & 3;
  } else {
    layer = 3;
    to = (tile - 40) >> 2;
    ti = (tile - 40) & 3;
  }
}

__global__ __launch_bounds__(64) void k_guide_wgrad(uint32_t count, const uint16_t* __restrict__ xs,
                                                    const uint16_t* __restrict__ ds, float4* __restrict__ part,
                                                    float* __restrict__ bpart) {
  const int tile = blockIdx.x, split = blockIdx.y, nsplit = gridDim.y;
  const int lane = threadIdx.x, r = lane & 15, gq = lane >> 4;
  int layer, to, ti;
  wtile(tile, layer, to, ti);
  const int xo[4] = {XO0, XO1, XO2, XO3};
  const int dof[4] = {DO1, DO2, DO3, DO4};
  const uint32_t pairs = count / 32u;
  const uint32_t p0 = (uint32_t)((uint64_t)pairs * split / nsplit), p1 = (uint32_t)((uint64_t)pairs * (split + 1) / nsplit);
  f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
  float bsum = 0.0f;
  const int half = (gq & 1) * 8;
  for (uint32_t pr = p0; pr < p1; ++pr) {
    size_t grp = 2ull * pr + (gq >> 1);
    uint4 a = *reinterpret_cast<const uint4*>(ds + grp * DS_GROUP + dof[layer] + (16 * to + r) * 16 + half);
    uint4 b = *reinterpret_cast<const uint4*>(xs + grp * XS_GROUP + xo[layer] + (16 * ti + r) * 16 + half);
    acc = mfma16(a, b, acc);
    if (ti == 0) {
      bsum += bf16_value(a.x & 0xffffu) + bf16_value(a.x >> 16) + bf16_value(a.y & 0xffffu) + bf16_value(a.y >> 16) +
              bf16_value(a.z & 0xffffu) + bf16_value(a.z >> 16) + bf16_value(a.w & 0xffffu) + bf16_value(a.w >> 16);
    }
  }
  size_t o = ((size_t)split * N_WTILES + tile) * 64 + lane;
  part[o] = make_float4(acc[0], acc[1], acc[2], acc[3]);
  bpart[o] = bsum;
}

// Sum of the split partials in split order into the MLP gradient.
__global__ void k_guide_wreduce(uint32_t nsplit, const float4* __restrict__ part, const float* __restrict__ bpart,
                                float* __restrict__ gmlp) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (uint32_t)G_NMLP) return;
  const int offs[4] = {G_OW1, G_OW2, G_OW3, G_OW4};
  const int boffs[4] = {G_OB1, G_OB2, G_OB3, G_OB4};
  const int nin[4] = {G_NIN, G_HID, G_HID, G_HID};
  const int tbase[4] = {0, 8, 24, 40};
  const int tper[4] = {2, 4, 4, 4};
  int layer = 3;
  for (int L = 0; L < 4; ++L)
    if ((int)k < boffs[L] + (L == 3 ? G_OUT : G_HID) && (int)k >= offs[L]) {
      layer = L;
      break;
    }
  float s = 0.0f;
  if ((int)k < boffs[layer]) {
    int e = (int)k - offs[layer];
    int o = e / nin[layer], i = e % nin[layer];
    int tile = tbase[layer] + (o >> 4) * tper[layer] + (i >> 4);
    int lane = (i & 15) + 16 * ((o & 15) >> 2);
    int j = o & 3;
    for (uint32_t sp = 0; sp < nsplit; ++sp) {
      float4 v = part[((size_t)sp * N_WTILES + tile) * 64 + lane];
      s += j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
    }
  } else {
    int o = (int)k - boffs[layer];
    int tile = tbase[layer] + (o >> 4) * tper[layer];
    for (uint32_t sp = 0; sp < nsplit; ++sp)
      for (int gq = 0; gq < 4; ++gq) s += bpart[((size_t)sp * N_WTILES + tile) * 64 + (o & 15) + 16 * gq];
  }
  gmlp[k] = s;
}

// Sums the gradient copies into ggrid and clears them. Float i of a copy
// is float i - off[l] of level l's dense table for the last l with
// off[l] <= i.
__global__ void k_guide_fold(uint32_t tmask, float* __restrict__ ggrid, GradCopies gc) {
  const size_t T = (size_t)tmask + 1u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < gc.stride; i += gridDim.x * blockDim.x) {
    uint32_t l = 0;
#pragma unroll
    for (uint32_t k = 1; k < G_COPY_LEVELS; ++k)
      if (k < gc.nlev && i >= gc.off[k]) l = k;
    float sum = 0.0f;
#pragma unroll
    for (int c = 0; c < G_GRAD_COPIES; ++c) {
      float* e = gc.p + (size_t)c * gc.stride + i;
      const float v = *e;
      if (v != 0.0f) {
        sum += v;
        *e = 0.0f;
      }
    }
    if (sum != 0.0f) ggrid[(size_t)l * T * 2 + (i - gc.off[l])] += sum;
  }
}

// Adam over [grid | mlp]; the grid part is lazy (zero gradient: no
// update) and its gradient is cleared for the next step's atomics.
__global__ void k_guide_adam(uint64_t n_grid, uint64_t n_total, float* __restrict__ p, float* __restrict__ m,
                             float* __restrict__ v, float* __restrict__ grad, float lr, float b1, float b2, float eps,
                             float c1, float c2) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n_total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    float gi = grad[i];
    if (i < n_grid) {
      if (gi == 0.0f) continue;
      grad[i] = 0.0f;
    }
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float mh = mi / c1, vh = vi / c2;
    p[i] -= lr * mh / (sqrtf(vh) + eps);
  }
}

// bf16 fragments of the weights in the MFMA K order (one block per
// fragment, one lane per thread).
OXR_D float frag_elem(const float* __restrict__ mlp, int f, int lane, int jp) {
  int r = lane & 15, g = lane >> 4;
  int kh = 16 * (jp >> 2) + 4 * g + (jp & 3);  // hidden K slot -> feature (within a 32-chunk)
  if (f < FW2) return mlp[G_OW1 + (16 * f + r) * G_NIN + 8 * g + jp];
  if (f < FW3) {
    int ff = f - FW2, o = 16 * (ff >> 1) + r, k = 32 * (ff & 1) + kh;
    return mlp[G_OW2 + o * G_HID + k];
  }
  if (f < FW4) {
    int ff = f - FW3, o = 16 * (ff >> 1) + r, k = 32 * (ff & 1) + kh;
    return mlp[G_OW3 + o * G_HID + k];
  }
  if (f < BW4) {
    int ff = f - FW4, o = 16 * (ff >> 1) + r, k = 32 * (ff & 1) + kh;
    return mlp[G_OW4 + o * G_HID + k];
  }
  if (f < BW3) {
    int ff = f - BW4, i = 16 * (ff >> 1) + r, o = 32 * (ff & 1) + kh;
    return o < G_OUT ? mlp[G_OW4 + o * G_HID + i] : 0.0f;
  }
  if (f < BW2) {
    int ff = f - BW3, i = 16 * (ff >> 1) + r, o = 32 * (ff & 1) + kh;
    return mlp[G_OW3 + o * G_HID + i];
  }
  if (f < BW1) {
    int ff = f - BW2, i = 16 * (ff >> 1) + r, o = 32 * (ff & 1) + kh;
    return mlp[G_OW2 + o * G_HID + i];
  }
  int ff = f - BW1, i = 8 * (r >> 2) + 4 * (ff >> 1) + (r & 3), o = 32 * (ff & 1) + kh;
  return mlp[G_OW1 + o * G_NIN + i];
}

__global__ __launch_bounds__(64) void k_guide_pack(const float* __restrict__ mlp, uint4* __restrict__ frag) {
  int f = blockIdx.x, lane = threadIdx.x;
  float v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = frag_elem(mlp, f, lane, j);
  frag[f * 64 + lane] = pack8(v);
}

// Strided batch of the collected samples: dst[i] = src[start + i * stride]
// for i < count (start < stride, start + (count - 1) * stride < the source
// count), so that every training step sees samples from the whole pass.
__global__ void k_guide_gather(const float4* __restrict__ src, uint32_t start, uint32_t stride, uint32_t count,
                               float4* __restrict__ dst) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  size_t j = (size_t)start + (size_t)i * stride;
  dst[2ull * i] = src[2ull * j];
  dst[2ull * i + 1] = src[2ull * j + 1];
}

__global__ void k_guide_fill(float* p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// Mixture pdf at host directions / samples from the mixture (tests).
__global__ void k_guide_pdf_sample(uint32_t n, const GuideLobes* __restrict__ lobes, const float* __restrict__ dir,
                                   const float* __restrict__ u, float* __restrict__ dir_out,
                                   float* __restrict__ pdf_out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const GuideLobes* L = lobes + i;
  V3 w;
  if (u) {
    w = guide_sample(L, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
    dir_out[3 * i] = w.x;
    dir_out[3 * i + 1] = w.y;
    dir_out[3 * i + 2] = w.z;
  } else {
    w = ld3(dir + 3 * i);
  }
  pdf_out[i] = guide_pdf(L, w);
}

// ---------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------
namespace glaunch {

static int cu_count() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  }
  return cus;
}

void infer(hipStream_t s, const DevGuideNet& n, uint32_t fixed_count, const uint32_t* count_ptr, uint32_t cap,
           const float* pos, const float4* qray, const uint4* hits, GuideLobes* out, float* raw) {
  uint32_t groups = (cap + 15u) / 16u;
  int grid = (int)std::min<uint32_t>((groups + 3u) / 4u, (uint32_t)cu_count() * 8u);
  if (grid < 1) grid = 1;
  if (pos)
    k_guide_infer<0><<<grid, GUIDE_TB, 0, s>>>(n, fixed_count, count_ptr, pos, qray, hits, out, raw);
  else
    k_guide_infer<1><<<grid, GUIDE_TB, 0, s>>>(n, fixed_count, count_ptr, pos, qray, hits, out, raw);
}

void train_grad(hipStream_t s, const DevGuideNet& n, uint32_t count, const float4* samp, float* inv_wsum,
                float* ggrid, const GradCopies& gc, float* gmlp, uint16_t* xs, uint16_t* ds, float4* part,
                float* bpart, uint32_t nsplit, float* loss) {
  k_guide_wsum<<<1, 1024, 0, s>>>(count, samp, inv_wsum, loss);
  uint32_t groups = count / 16u;
  int grid = (int)std::min<uint32_t>((groups + 3u) / 4u, (uint32_t)cu_count() * 4u);
  if (grid < 1) grid = 1;
  k_guide_fwdbwd<<<grid, GUIDE_TB, 0, s>>>(n, count, samp, inv_wsum, ggrid, gc, xs, ds, loss);
  if (gc.nlev) {
    int fgrid = (int)std::min<uint32_t>((gc.stride + 255u) / 256u, (uint32_t)cu_count() * 4u);
    k_guide_fold<<<fgrid, 256, 0, s>>>(n.tmask, ggrid, gc);
  }
  k_guide_wgrad<<<dim3(N_WTILES, nsplit), 64, 0, s>>>(count, xs, ds, part, bpart);
  k_guide_wreduce<<<(G_NMLP + 255) / 256, 256, 0, s>>>(nsplit, part, bpart, gmlp);
}

void adam(hipStream_t s, uint64_t n_grid, uint64_t n_total, float* p, float* m, float* v, float* grad, float lr,
          float b1, float b2, float eps, float c1, float c2) {
  int grid = (int)std::min<uint64_t>((n_total + 255) / 256, (uint64_t)cu_count() * 16u);
  k_guide_adam<<<grid, 256, 0, s>>>(n_grid, n_total, p, m, v, grad, lr, b1, b2, eps, c1, c2);
}

void pack(hipStream_t s, const float* mlp, uint4* frag) { k_guide_pack<<<NFRAG, 64, 0, s>>>(mlp, frag); }

void fill(hipStream_t s, float* p, uint64_t n) {
  int grid = (int)std::min<uint64_t>((n + 255) / 256, (uint64_t)cu_count() * 16u);
  if (grid < 1) grid = 1;
  k_guide_fill<<<grid, 256, 0, s>>>(p, n);
}

void gather(hipStream_t s, const float4* src, uint32_t start, uint32_t stride, uint32_t count, float4* dst) {
  if (count == 0) return;
  k_guide_gather<<<(count + 255) / 256, 256, 0, s>>>(src, start, stride, count, dst);
}

void pdf_sample(hipStream_t s, uint32_t n, const GuideLobes* lobes, const float* dir, const float* u, float* dir_out,
                float* pdf_out) {
  k_guide_pdf_sample<<<(n + 255) / 256, 256, 0, s>>>(n, lobes, dir, u, dir_out, pdf_out);
}

uint32_t wtiles() { return N_WTILES; }
size_t xs_group() { return XS_GROUP; }
size_t ds_group() { return DS_GROUP; }

}  // namespace glaunch

}  // namespace oxr
