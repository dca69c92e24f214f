// oxr-cpu guide — the ORACLE of the learned path guide (test
// infrastructure, linked into liboxr_cpu.so; the product never calls it).
//
// oxideav-render has no path guiding, so there is no reference file to
// follow. This is the CPU statement of the design in include/oxr/guide.h:
//  - the hash-grid encoding;
//  - the 32-64-64-64-48 MLP;
//  - the vMF-mixture decoding, pdf and sampling;
//  - the cross-entropy loss and its analytic gradient;
//  - the Adam step.
// The device kernels must agree with it.
//
// Two arithmetic modes:
//   mode 0  exact: everything in f64, no rounding of operands. The
//           finite-difference gradient check pins this mode.
//   mode 1  device: the device's precision choices.
//           - bf16 (round-to-nearest-even) matmul operands: the encoded
//             features, the weights, the ReLU outputs and the back-propagated
//             deltas.
//           - f32 accumulation starting from the bias.
//           - f32 lobe math.
//
// Pins (tests/test_oracle_guide.py):
//  - central finite differences of the mode-0 loss against its analytic
//    gradient;
//  - trilinear reproduction of a linear field on dense levels;
//  - the vMF mixture integrating to 1 over the sphere;
//  - sampling histograms against the pdf;
//  - mode 1 against mode 0 within bf16 precision.

#include "oxr/guide.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int L = OXR_GUIDE_LEVELS;
constexpr int H = OXR_GUIDE_HIDDEN;
constexpr int NO = OXR_GUIDE_OUTPUTS;
constexpr int NL = OXR_GUIDE_LOBES;
constexpr int NIN = L * OXR_GUIDE_FEATURES;

// Offsets of the MLP part of the parameter array.
constexpr int OW1 = 0, OB1 = OW1 + H * NIN, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H * H,
              OW4 = OB3 + H, OB4 = OW4 + NO * H, NMLP = OB4 + NO;
static_assert(NMLP == OXR_GUIDE_MLP_PARAMS, "MLP parameter count");

float bf16r(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return x;  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

struct Grid {
  uint32_t res[L];
  uint32_t T;
  bool dense[L];
  float lo[3], inv[3];
};

Grid make_grid(const oxr_guide_config* c) {
  Grid g;
  g.T = 1u << c->log2_table;
  double b = std::exp((std::log((double)c->max_res) - std::log((double)c->base_res)) / (L - 1));
  for (int l = 0; l < L; ++l) {
    g.res[l] = (uint32_t)std::floor((double)c->base_res * std::pow(b, l) + 1e-6);
    double v = (double)g.res[l] + 1.0;
    g.dense[l] = v * v * v <= (double)g.T;
  }
  for (int d = 0; d < 3; ++d) {
    g.lo[d] = c->aabb_min[d];
    float e = c->aabb_max[d] - c->aabb_min[d];
    g.inv[d] = e > 0.0f ? 1.0f / e : 0.0f;
  }
  return g;
}

uint32_t grid_index(const Grid& g, int l, uint32_t x, uint32_t y, uint32_t z) {
  if (g.dense[l]) {
    uint64_t r = (uint64_t)g.res[l] + 1;
    return (uint32_t)(x + r * (y + r * z));
  }
  return ((x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u)) & (g.T - 1u);
}

// Corner indices and trilinear weights of one level (f32, device order:
// corner bit 0 = x, bit 1 = y, bit 2 = z; weight = wx * wy * wz).
void corners(const Grid& g, int l, const float* p, uint32_t idx[8], float w[8]) {
  float fr[3];
  uint32_t base[3];
  float R = (float)g.res[l];
  for (int d = 0; d < 3; ++d) {
    float x = (p[d] - g.lo[d]) * g.inv[d];
    x = x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f;  // NaN -> 0
    float s = x * R;
    float fl = std::floor(s);
    int i = (int)fl;
    if (i > (int)g.res[l] - 1) i = (int)g.res[l] - 1;
    base[d] = (uint32_t)i;
    fr[d] = s - (float)i;
  }
  for (int c = 0; c < 8; ++c) {
    uint32_t cx = base[0] + (c & 1), cy = base[1] + ((c >> 1) & 1), cz = base[2] + ((c >> 2) & 1);
    idx[c] = grid_index(g, l, cx, cy, cz);
    float wx = (c & 1) ? fr[0] : 1.0f - fr[0];
    float wy = (c & 2) ? fr[1] : 1.0f - fr[1];
    float wz = (c & 4) ? fr[2] : 1.0f - fr[2];
    w[c] = wx * wy * wz;
  }
}

// Output index of (lobe, component): lane group g = lobe / 2 owns lobes
// 2g, 2g+1 at local slots e = 5 * (lobe % 2) + comp, slot e at tile e / 4,
// row 4g + e % 4.
int out_index(int lobe, int comp) {
  int g = lobe / 2, e = 5 * (lobe % 2) + comp;
  return 16 * (e / 4) + 4 * g + (e % 4);
}

template <class T>
struct Net {
  T x0[NIN], z1[H], x1[H], z2[H], x2[H], z3[H], x3[H], z4[NO];
};

// Forward of one sample. mode 1 rounds operands to bf16 and accumulates in
// f32; mode 0 uses T = double throughout.
template <class T>
void forward(const Grid& g, const float* grid, const float* mlp, const float* p, int mode, Net<T>& n,
             uint32_t idx[L][8], float wts[L][8]) {
  for (int l = 0; l < L; ++l) {
    corners(g, l, p, idx[l], wts[l]);
    for (int f = 0; f < 2; ++f) {
      T acc = 0;
      for (int c = 0; c < 8; ++c) acc += (T)wts[l][c] * (T)grid[((size_t)l * g.T + idx[l][c]) * 2 + f];
      n.x0[2 * l + f] = mode ? (T)bf16r((float)acc) : acc;
    }
  }
  auto layer = [&](const T* x, int nin, int ow, int ob, int nout, T* z, T* xo) {
    for (int o = 0; o < nout; ++o) {
      T acc = (T)mlp[ob + o];
      for (int k = 0; k < nin; ++k) {
        T w = mode ? (T)bf16r(mlp[ow + o * nin + k]) : (T)mlp[ow + o * nin + k];
        acc += w * x[k];
      }
      z[o] = acc;
      if (xo) {
        T r = acc > 0 ? acc : (T)0;
        xo[o] = mode ? (T)bf16r((float)r) : r;
      }
    }
  };
  layer(n.x0, NIN, OW1, OB1, H, n.z1, n.x1);
  layer(n.x1, H, OW2, OB2, H, n.z2, n.x2);
  layer(n.x2, H, OW3, OB3, H, n.z3, n.x3);
  layer(n.x3, H, OW4, OB4, NO, n.z4, nullptr);
}

template <class T>
struct Lobes {
  T mu[NL][3], kappa[NL], pi[NL], logpi[NL], len[NL], s[NL];
};

template <class T>
void decode(const T* raw, Lobes<T>& lb) {
  T mx = -INFINITY;
  for (int k = 0; k < NL; ++k) mx = std::max(mx, raw[out_index(k, 4)]);
  T sum = 0;
  T e[NL];
  for (int k = 0; k < NL; ++k) {
    e[k] = std::exp(raw[out_index(k, 4)] - mx);
    sum += e[k];
  }
  T lsum = std::log(sum);
  for (int k = 0; k < NL; ++k) {
    T m0 = raw[out_index(k, 0)], m1 = raw[out_index(k, 1)], m2 = raw[out_index(k, 2)];
    T len = std::sqrt(m0 * m0 + m1 * m1 + m2 * m2);
    lb.len[k] = len;
    if (len > (T)1e-12) {
      lb.mu[k][0] = m0 / len;
      lb.mu[k][1] = m1 / len;
      lb.mu[k][2] = m2 / len;
    } else {
      lb.mu[k][0] = 0;
      lb.mu[k][1] = 0;
      lb.mu[k][2] = 1;
    }
    T s = raw[out_index(k, 3)];
    lb.s[k] = s;
    T sc = s < (T)-3 ? (T)-3 : (s > (T)8 ? (T)8 : s);
    lb.kappa[k] = std::exp(sc);
    lb.pi[k] = e[k] / sum;
    lb.logpi[k] = raw[out_index(k, 4)] - mx - lsum;
  }
}

// log of the vMF density at cosine c = mu . w.
template <class T>
T log_vmf(T kappa, T c) {
  return std::log(kappa) - std::log((T)(2.0 * M_PI)) - std::log(-std::expm1(-2 * kappa)) + kappa * (c - 1);
}

template <class T>
T dot3(const T* a, const float* b) {
  return a[0] * (T)b[0] + a[1] * (T)b[1] + a[2] * (T)b[2];
}

template <class T>
T loss_grad_one(const Grid& g, const float* grid, const float* mlp, const float* p, const float* dir, T wn, int mode,
                double* gg, double* gm) {
  Net<T> n;
  uint32_t idx[L][8];
  float wts[L][8];
  forward<T>(g, grid, mlp, p, mode, n, idx, wts);
  Lobes<T> lb;
  decode<T>(n.z4, lb);
  T lp[NL];
  T mx = -INFINITY;
  for (int k = 0; k < NL; ++k) {
    lp[k] = lb.logpi[k] + log_vmf<T>(lb.kappa[k], dot3<T>(lb.mu[k], dir));
    mx = std::max(mx, lp[k]);
  }
  T se = 0;
  for (int k = 0; k < NL; ++k) se += std::exp(lp[k] - mx);
  T lse = mx + std::log(se);
  T loss = -wn * lse;
  if (!gg && !gm) return loss;
  // d loss / d raw.
  T d4[NO];
  for (int o = 0; o < NO; ++o) d4[o] = 0;
  for (int k = 0; k < NL; ++k) {
    T r = std::exp(lp[k] - lse);
    T kap = lb.kappa[k];
    T c = dot3<T>(lb.mu[k], dir);
    d4[out_index(k, 4)] = -wn * (r - lb.pi[k]);
    if (lb.s[k] > (T)-3 && lb.s[k] < (T)8)
      d4[out_index(k, 3)] = -wn * r * (1 - 2 * kap / std::expm1(2 * kap) + kap * (c - 1));
    if (lb.len[k] > (T)1e-12) {
      for (int d = 0; d < 3; ++d)
        d4[out_index(k, d)] = -wn * r * kap * ((T)dir[d] - c * lb.mu[k][d]) / lb.len[k];
    }
  }
  auto rd = [&](T v) -> T { return mode ? (T)bf16r((float)v) : v; };
  auto wv = [&](int off) -> T { return mode ? (T)bf16r(mlp[off]) : (T)mlp[off]; };
  // Back through one layer: dz (rounded) -> weight/bias grads, dx = W^T dz.
  auto back = [&](const T* dz, const T* x, int nin, int ow, int ob, int nout, T* dx) {
    for (int o = 0; o < nout; ++o) {
      if (gm) {
        gm[ob + o] += (double)dz[o];
        for (int k = 0; k < nin; ++k) gm[ow + o * nin + k] += (double)dz[o] * (double)x[k];
      }
    }
    for (int k = 0; k < nin; ++k) {
      T acc = 0;
      for (int o = 0; o < nout; ++o) acc += wv(ow + o * nin + k) * dz[o];
      dx[k] = acc;
    }
  };
  T dz4[NO], dx3[H], dz3[H], dx2[H], dz2[H], dx1[H], dz1[H], dx0[NIN];
  for (int o = 0; o < NO; ++o) dz4[o] = rd(d4[o]);
  back(dz4, n.x3, H, OW4, OB4, NO, dx3);
  for (int o = 0; o < H; ++o) dz3[o] = rd(n.z3[o] > 0 ? dx3[o] : (T)0);
  back(dz3, n.x2, H, OW3, OB3, H, dx2);
  for (int o = 0; o < H; ++o) dz2[o] = rd(n.z2[o] > 0 ? dx2[o] : (T)0);
  back(dz2, n.x1, H, OW2, OB2, H, dx1);
  for (int o = 0; o < H; ++o) dz1[o] = rd(n.z1[o] > 0 ? dx1[o] : (T)0);
  back(dz1, n.x0, NIN, OW1, OB1, H, dx0);
  if (gg) {
    for (int l = 0; l < L; ++l)
      for (int c = 0; c < 8; ++c)
        for (int f = 0; f < 2; ++f)
          gg[((size_t)l * g.T + idx[l][c]) * 2 + f] += (double)wts[l][c] * (double)dx0[2 * l + f];
  }
  return loss;
}

}  // namespace

extern "C" {

void oxrcpu_guide_resolutions(const oxr_guide_config* cfg, uint32_t* res, uint8_t* dense) {
  Grid g = make_grid(cfg);
  for (int l = 0; l < L; ++l) {
    res[l] = g.res[l];
    if (dense) dense[l] = g.dense[l];
  }
}

float oxrcpu_bf16_round(float x) { return bf16r(x); }

// Per level: the 8 corner indices and weights of each position (n x L x 8).
void oxrcpu_guide_corners(const oxr_guide_config* cfg, uint32_t n, const float* pos, uint32_t* idx, float* w) {
  Grid g = make_grid(cfg);
  for (uint32_t i = 0; i < n; ++i)
    for (int l = 0; l < L; ++l) corners(g, l, pos + 3 * i, idx + ((size_t)i * L + l) * 8, w + ((size_t)i * L + l) * 8);
}

// Encoded features (n x 32) and raw outputs (n x 48) of the network.
void oxrcpu_guide_forward(const oxr_guide_config* cfg, const float* grid, const float* mlp, uint32_t n,
                          const float* pos, int mode, float* feat, float* raw) {
  Grid g = make_grid(cfg);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t idx[L][8];
    float wts[L][8];
    if (mode) {
      Net<float> net;
      forward<float>(g, grid, mlp, pos + 3 * i, 1, net, idx, wts);
      if (feat) std::copy(net.x0, net.x0 + NIN, feat + (size_t)i * NIN);
      if (raw) std::copy(net.z4, net.z4 + NO, raw + (size_t)i * NO);
    } else {
      Net<double> net;
      forward<double>(g, grid, mlp, pos + 3 * i, 0, net, idx, wts);
      for (int k = 0; k < NIN && feat; ++k) feat[(size_t)i * NIN + k] = (float)net.x0[k];
      for (int k = 0; k < NO && raw; ++k) raw[(size_t)i * NO + k] = (float)net.z4[k];
    }
  }
}

// Lobes (n x 8 x 5: mu.xyz, kappa, pi) from raw outputs (f32 math).
void oxrcpu_guide_lobes(uint32_t n, const float* raw, float* lobes) {
  for (uint32_t i = 0; i < n; ++i) {
    Lobes<float> lb;
    decode<float>(raw + (size_t)i * NO, lb);
    for (int k = 0; k < NL; ++k) {
      float* o = lobes + ((size_t)i * NL + k) * 5;
      o[0] = lb.mu[k][0];
      o[1] = lb.mu[k][1];
      o[2] = lb.mu[k][2];
      o[3] = lb.kappa[k];
      o[4] = lb.pi[k];
    }
  }
}

// Mixture pdf (f64 evaluation of the f32 lobes) at unit directions.
void oxrcpu_guide_pdf(uint32_t n, const float* lobes, const float* dir, double* pdf) {
  for (uint32_t i = 0; i < n; ++i) {
    double q = 0;
    for (int k = 0; k < NL; ++k) {
      const float* o = lobes + ((size_t)i * NL + k) * 5;
      double c = (double)o[0] * dir[3 * i] + (double)o[1] * dir[3 * i + 1] + (double)o[2] * dir[3 * i + 2];
      q += (double)o[4] * std::exp(log_vmf<double>((double)o[3], c));
    }
    pdf[i] = q;
  }
}

// Mixture sampling with u = (lobe, u1, u2), f32 arithmetic in the device's
// order: lobe by the running sum of pi, w = 1 + log1p(-(1-u1)(1-e^{-2k}))/k,
// phi = 2 pi u2 in the (t, b) frame of mu (Duff et al. 2017).
void oxrcpu_guide_sample(uint32_t n, const float* lobes, const float* u, float* dir) {
  for (uint32_t i = 0; i < n; ++i) {
    const float* lb = lobes + (size_t)i * NL * 5;
    int pick = NL - 1;
    float c = 0.0f;
    for (int k = 0; k < NL; ++k) {
      c += lb[5 * k + 4];
      if (u[3 * i] < c) {
        pick = k;
        break;
      }
    }
    const float* o = lb + 5 * pick;
    float kap = o[3];
    float u1 = u[3 * i + 1], u2 = u[3 * i + 2];
    float w = 1.0f + std::log1p(-(1.0f - u1) * -std::expm1(-2.0f * kap)) / kap;
    w = w < -1.0f ? -1.0f : (w > 1.0f ? 1.0f : w);
    float st = std::sqrt(std::max(0.0f, 1.0f - w * w));
    float phi = 2.0f * (float)M_PI * u2;
    float nx = o[0], ny = o[1], nz = o[2];
    float sign = std::copysign(1.0f, nz);
    float a = -1.0f / (sign + nz);
    float bb = nx * ny * a;
    float t[3] = {1.0f + sign * nx * nx * a, sign * bb, -sign * nx};
    float b[3] = {bb, sign + ny * ny * a, -ny};
    float cx = st * std::cos(phi), cy = st * std::sin(phi);
    for (int d = 0; d < 3; ++d) dir[3 * i + d] = t[d] * cx + b[d] * cy + o[d] * w;
  }
}

// Loss sum_i w_i * -log q_i / sum_i w_i and its gradient (added into
// grad_grid / grad_mlp, f64, either may be null).
double oxrcpu_guide_loss_grad(const oxr_guide_config* cfg, const float* grid, const float* mlp, uint32_t n,
                              const float* pos, const float* dir, const float* weight, int mode, double* grad_grid,
                              double* grad_mlp) {
  Grid g = make_grid(cfg);
  double sw = 0;
  for (uint32_t i = 0; i < n; ++i) sw += weight[i];
  if (!(sw > 0)) return 0.0;
  double loss = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (mode) {
      float wn = (float)((double)weight[i] / sw);
      loss += loss_grad_one<float>(g, grid, mlp, pos + 3 * i, dir + 3 * i, wn, 1, grad_grid, grad_mlp);
    } else {
      double wn = (double)weight[i] / sw;
      loss += loss_grad_one<double>(g, grid, mlp, pos + 3 * i, dir + 3 * i, wn, 0, grad_grid, grad_mlp);
    }
  }
  return loss;
}

// One Adam step (t = step number, 1-based). lazy != 0 skips entries whose
// gradient is exactly zero (the hash grid).
void oxrcpu_guide_adam(uint64_t n, float* p, float* m, float* v, const float* grad, uint32_t t, float lr, float b1,
                       float b2, float eps, int lazy) {
  float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
  for (uint64_t i = 0; i < n; ++i) {
    float gi = grad[i];
    if (lazy && gi == 0.0f) continue;
    m[i] = b1 * m[i] + (1.0f - b1) * gi;
    v[i] = b2 * v[i] + (1.0f - b2) * gi * gi;
    float mh = m[i] / c1, vh = v[i] / c2;
    p[i] -= lr * mh / (std::sqrt(vh) + eps);
  }
}

}  // extern "C"
