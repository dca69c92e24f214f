// HIP kernels of the oxr wavefront path tracer (gfx950).
//
// One refine batch = P paths (NPX pixels x S samples). Per bounce:
//   k_prep     reset the next queue / shadow queue / work counters
//   k_closest  persistent closest-hit traversal of the ray queue
//   k_shade    hit -> emission, material, BSDF, NEE shadow rays, BSDF
//              sample, roulette; appends next rays and shadow rays
//   k_anyhit   persistent occlusion traversal of the shadow queue
// then k_finalize (pending NEE, per-sample radiance) and k_reduce
// (per-pixel f64 accumulation in sample order).
//
// The estimator is oxideav-render's (src/pathtrace.rs:2004-2329,
// normative spec :1-277); each device function cites its counterpart.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bvh.h"
#include "dev_bsdf.h"
#include "dev_guide.h"
#include "dev_math.h"
#include "dev_scene.h"
#include "runtime.h"

namespace oxr {

__constant__ uint32_t c_sobol_tab[3][4][256];
__constant__ float c_sheen[16 * 32];
__constant__ float c_srgb_th[255];

constexpr int TRACE_TB = 128;
// LDS traversal stack entries per lane (a ring; deeper entries spill to the
// global stack). 16 entries = 8 KB per block leaves the traversal kernels
// limited by registers (7 waves per SIMD), not LDS; 32 entries capped them
// at 5 by LDS and measured 2-4% slower (material sweep, terrain).
constexpr int STACK = 16;
constexpr float SLAB_EPS = 1.0f + 1.0f / 65536.0f;

// Reciprocal of a direction component for the slab test. A component below
// 2^-64 in magnitude (zero included: axis-parallel rays occur on flat
// ground and mirror bounces) takes the finite reciprocal +-2^64. With an
// infinite one the per-node offset org * inv - o * inv is inf - inf = NaN,
// fmaxf / fminf drop the NaN, and the axis culls nothing: the ray visits
// every node its other two slabs overlap (8330 nodes and 15589 triangles
// for one such ray on a 12.4M-triangle heightfield, against 37 nodes on
// average), holding its wave for milliseconds. The decoded boxes hold the
// geometry with a margin of one quantisation step, far above the
// cancellation error of the finite form, so no box the ray passes through
// is culled.
OXR_D float slab_rcp(float d) { return fabsf(d) < 0x1p-64f ? copysignf(0x1p64f, d) : 1.0f / d; }

OXR_D uint32_t lane_id() { return __lane_id(); }

// Wave-aggregated append: one atomic per wave; returns this lane's slot.
OXR_D uint32_t wave_append(uint32_t* counter, bool pred) {
  uint64_t active = __ballot(1);
  uint64_t m = __ballot(pred);
  uint32_t total = (uint32_t)__popcll(m);
  int leader = __ffsll((unsigned long long)active) - 1;
  uint32_t lane = lane_id();
  uint32_t base = 0;
  if ((int)lane == leader && total) base = atomicAdd(counter, total);
  base = __shfl(base, leader);
  uint64_t below = lane == 0 ? 0ull : (m & ((~0ull) >> (64 - lane)));
  return base + (uint32_t)__popcll(below);
}

// Queue staging of the shading kernel. Appends to one queue counter are
// device-scope atomics that serialise device-wide, so a block gathers its
// round's queue entries in LDS and appends them with one atomic per queue.
// stage_entry: an entering lane stores its entry at the wave's next free
// staging entries (all 64 lanes present, `staged` wave-uniform).
OXR_D void stage_entry(float4* sa, float4* sb, uint32_t* si, uint32_t wbase, uint32_t& staged, bool enter,
                       float4 a, float4 b, uint32_t idx) {
  const uint64_t m = __ballot(enter);
  const uint32_t lane = lane_id();
  if (enter) {
    const uint64_t below = lane == 0 ? 0ull : (m & ((~0ull) >> (64 - lane)));
    const uint32_t e = wbase + staged + (uint32_t)__popcll(below);
    sa[e] = a;
    sb[e] = b;
    si[e] = idx;
  }
  staged += (uint32_t)__popcll(m);
}

// The block's staged entries of two queues (shadow rays and next rays) go
// out with one atomic per queue: every thread of the block calls this once
// per round. s_cnt / s_base: 8 words each.
OXR_D void block_flush(uint32_t* s_cnt, uint32_t* s_base, const float4* sa, const float4* sb, const uint32_t* si,
                       uint32_t sw, uint32_t& sstaged, uint32_t* scount, float4* __restrict__ sray,
                       uint32_t* __restrict__ sidx, const float4* na, const float4* nb, const uint32_t* np,
                       uint32_t nw, uint32_t& nstaged, uint32_t* ncount, float4* __restrict__ nray,
                       uint32_t* __restrict__ npid) {
  const uint32_t w = threadIdx.x / 64u, lane = lane_id();
  if (lane == 0u) {
    s_cnt[w] = sstaged;
    s_cnt[4 + w] = nstaged;
  }
  __syncthreads();
  if (lane == 0u && w < 2u) {  // wave 0: shadow queue, wave 1: next-ray queue
    const uint32_t* c = s_cnt + 4u * w;
    const uint32_t c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
    const uint32_t tot = c0 + c1 + c2 + c3;
    const uint32_t b = tot ? atomicAdd(w == 0u ? scount : ncount, tot) : 0u;
    uint32_t* o = s_base + 4u * w;
    o[0] = b;
    o[1] = b + c0;
    o[2] = b + c0 + c1;
    o[3] = b + c0 + c1 + c2;
  }
  __syncthreads();
  const uint32_t bs = s_base[w], bn = s_base[4 + w];
  for (uint32_t e = lane; e < sstaged; e += 64u) {
    const uint32_t q = bs + e;
    sray[2 * (size_t)q] = sa[sw + e];
    sray[2 * (size_t)q + 1] = sb[sw + e];
    sidx[q] = si[sw + e];
  }
  for (uint32_t e = lane; e < nstaged; e += 64u) {
    const uint32_t q = bn + e;
    nray[2 * (size_t)q] = na[nw + e];
    nray[2 * (size_t)q + 1] = nb[nw + e];
    npid[q] = np[nw + e];
  }
  sstaged = 0;
  nstaged = 0;
}

// ---------------------------------------------------------------------
// BVH traversal (TraceScene::closest_hit_filtered / occluded_filtered,
// src/trace.rs:405-450), 4-wide quantised nodes (Node4Q), per-thread
// stack: a STACK-entry ring in LDS with overflow spilled to a per-thread
// global buffer whose size the host derives from the tree depth.
// ---------------------------------------------------------------------
struct Stack {
  uint32_t* lds;    // &s[0][tid]; entry i at lds[i * TRACE_TB]
  uint32_t* spill;  // per-thread overflow
  uint32_t cap;
  uint32_t sp, bottom, spn;
  bool overflow;

  // Pushes the last k (0..3) of c0, c1, c2 in that order (the children to
  // visit sorted far to near are a suffix). Room is made first by moving
  // bottom entries to the global spill, which leaves the stack as k single
  // pushes would; no pushed entry can reach the bottom (k < STACK).
  OXR_D void push_suffix(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t k) {
    if (sp + k > STACK) {
      do {
        if (spn >= cap) {
          overflow = true;
          return;
        }
        spill[spn++] = lds[(bottom & (STACK - 1)) * TRACE_TB];
        bottom = (bottom + 1) & (STACK - 1);
        --sp;
      } while (sp + k > STACK);
    }
    const uint32_t e0 = k == 3u ? c0 : (k == 2u ? c1 : c2), e1 = k == 3u ? c1 : c2;
    if (k >= 1u) lds[((bottom + sp) & (STACK - 1)) * TRACE_TB] = e0;
    if (k >= 2u) lds[((bottom + sp + 1u) & (STACK - 1)) * TRACE_TB] = e1;
    if (k == 3u) lds[((bottom + sp + 2u) & (STACK - 1)) * TRACE_TB] = c2;
    sp += k;
  }
  OXR_D bool pop(uint32_t& x) {
    if (sp > 0) {
      --sp;
      x = lds[((bottom + sp) & (STACK - 1)) * TRACE_TB];
      return true;
    }
    if (spn > 0) {
      x = spill[--spn];
      return true;
    }
    return false;
  }
};

// Acceptance of a candidate (Integrator::accept, src/pathtrace.rs:1924-1934)
// from the per-triangle flags: primary back-face culling of single-sided
// materials, MASK cutoff, stochastic BLEND coin.
OXR_D bool accept_tri(uint32_t flags, float alpha, bool front, bool cull, uint32_t coin_key, uint32_t gid) {
  if (cull && !front && !(flags & TRI_DOUBLE_SIDED)) return false;
  if (flags & TRI_MASK_REJECT) return false;
  if (flags & TRI_BLEND) return to_unit(hash_mix(coin_key, gid)) < alpha;  // coin(path_seed, ray_id, gid)
  return true;
}

// One ray's traversal state. step() visits one node (its leaves are
// tested right away) and returns true once the ray is finished.
// PATH: rays of the wavefront (t_min 0, every triangle flag seen); else
// rays of the parity entry points.
template <bool ANY, bool ALPHA = false, bool PATH = true>
struct Trav {
  const DevScene* asc;  // ALPHA: the scene whose textured alpha the filter evaluates
  V3 inv, oi;
  WRay wr;
  float tmin, tmax;
  bool negx, negy, negz, cull;
  // skip: the global id an occlusion ray ignores. coin_key:
  // hash_mix(path_seed, ray_id), the ray's part of the BLEND coin
  // coin(path_seed, ray_id, gid) (one register instead of two, and no
  // per-leaf rehash).
  uint32_t skip, coin_key;
  // Triangle flags the filter sees: all of them on path rays; only the
  // double-sided bit on the parity entry points, which accept every
  // candidate (TraceScene::closest_hit / occluded, src/trace.rs:396-450)
  // apart from the optional back-face cull.
  static constexpr uint32_t fmask = PATH ? ~0u : TRI_DOUBLE_SIDED;
  uint32_t best;
  float bu, bv;
  uint32_t node;
  bool hit_any;

  OXR_D void setup(V3 o, V3 d, float tmin_, float tmax_, bool cull_, uint32_t skip_, uint32_t coin_key_, Stack& st) {
    inv = v3(slab_rcp(d.x), slab_rcp(d.y), slab_rcp(d.z));
    oi = mul(o, inv);
    negx = signbit(inv.x);
    negy = signbit(inv.y);
    negz = signbit(inv.z);
    wr = prepare_wray(o, d);
    tmin = PATH ? 0.0f : tmin_;
    tmax = tmax_;
    cull = cull_;
    skip = skip_;
    coin_key = coin_key_;
    best = HIT_MISS;
    bu = bv = 0.0f;
    node = 0;
    hit_any = false;
    st.sp = 0;
    st.bottom = 0;
    st.spn = 0;
    st.overflow = false;
  }

  // Textured alpha of a candidate (TraceScene::alpha, src/trace.rs:
  // 492-515): MASK compares it with the cutoff held in the triangle
  // record; BLEND's coin then uses it in place of the constant alpha.
  OXR_D bool tex_alpha_filter(uint32_t fl, uint32_t gid, float, float b1, float b2, float& aw) const {
    if (!(fl & TRI_ALPHA_TEX)) return true;
    uint32_t mid = asc->tri_info[gid] & TI_MAT_MASK;
    float al = tex_alpha(asc->tex, asc->mats[mid].base[3], mid, gid, v3(1.0f - b1 - b2, b1, b2));
    if ((fl & TRI_MASK_TEX) && !(al >= aw)) return false;
    aw = al;
    return true;
  }

  // nodes / tris: the scene's BVH (uniform across the wave when every
  // ray of the launch belongs to one scene, so they stay in SGPRs).
  // Node4Q: a plane q (byte) of axis k sits at org[k] + q * 2^e[k], so its
  // slab distance is q * (2^e[k] * inv[k]) + (org[k] * inv[k] - oi[k]).
  // COUNT: ntri counts the triangles fetched.
  template <bool COUNT = false>
  OXR_D bool step(Stack& st, const Node4Q* __restrict__ nodes, const TriRec* __restrict__ tris, uint32_t& ntri) {
    const uint4* np = reinterpret_cast<const uint4*>(nodes + node);
    const uint4 q0 = np[0], q1 = np[1], q2 = np[2], ch = np[3];
    const float sx = ldexpf(inv.x, (int)(int8_t)(q0.w & 0xffu));
    const float sy = ldexpf(inv.y, (int)(int8_t)((q0.w >> 8) & 0xffu));
    const float sz = ldexpf(inv.z, (int)(int8_t)((q0.w >> 16) & 0xffu));
    const float ox = fmaf(__uint_as_float(q0.x), inv.x, -oi.x);
    const float oy = fmaf(__uint_as_float(q0.y), inv.y, -oi.y);
    const float oz = fmaf(__uint_as_float(q0.z), inv.z, -oi.z);
    uint32_t chl[4] = {ch.x, ch.y, ch.z, ch.w};
    uint32_t cnt[4] = {q2.z & 0xffu, (q2.z >> 8) & 0xffu, (q2.z >> 16) & 0xffu, q2.z >> 24};
    // Near and far byte planes of all four children, chosen once per axis
    // by the ray's direction sign (a select on the packed words instead of
    // one per child and plane).
    const uint32_t pnx = negx ? q1.y : q1.x, pfx = negx ? q1.x : q1.y;
    const uint32_t pny = negy ? q1.w : q1.z, pfy = negy ? q1.z : q1.w;
    const uint32_t pnz = negz ? q2.y : q2.x, pfz = negz ? q2.x : q2.y;
    float tn[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t sh = 8u * c;
      float nx = fmaf((float)((pnx >> sh) & 0xffu), sx, ox), fx = fmaf((float)((pfx >> sh) & 0xffu), sx, ox);
      float ny = fmaf((float)((pny >> sh) & 0xffu), sy, oy), fy = fmaf((float)((pfy >> sh) & 0xffu), sy, oy);
      float nz = fmaf((float)((pnz >> sh) & 0xffu), sz, oz), fz = fmaf((float)((pfz >> sh) & 0xffu), sz, oz);
      float tnear = fmaxf(fmaxf(fmaxf(tmin, nx), ny), nz);
      float tfar = fminf(fminf(fminf(tmax, fx), fy), fz);
      bool hit = chl[c] != NODE_EMPTY && tnear <= tfar * SLAB_EPS;
      tn[c] = hit ? tnear : __builtin_huge_valf();
    }
    // The triangles of the hit leaf children, as one run in child order:
    // e0..e3 are the running ends. One loop over the run (rather than one
    // per child) lets a wave test lanes whose leaves sit under different
    // children together; each lane tests the same triangles in the same
    // order as child by child.
    uint32_t e0, e1, e2, e3;
    {
      const uint32_t n0 = (tn[0] != __builtin_huge_valf()) ? cnt[0] : 0u;
      const uint32_t n1 = (tn[1] != __builtin_huge_valf()) ? cnt[1] : 0u;
      const uint32_t n2 = (tn[2] != __builtin_huge_valf()) ? cnt[2] : 0u;
      const uint32_t n3 = (tn[3] != __builtin_huge_valf()) ? cnt[3] : 0u;
      e0 = n0;
      e1 = e0 + n1;
      e2 = e1 + n2;
      e3 = e2 + n3;
    }
    if (COUNT && !ANY) ntri += e3;
    for (uint32_t j = 0; j < e3; ++j) {
      if (COUNT && ANY) ++ntri;  // occlusion rays stop at the first accepted triangle
      const uint32_t ti = j < e0 ? chl[0] + j : (j < e1 ? chl[1] + (j - e0) : (j < e2 ? chl[2] + (j - e1) : chl[3] + (j - e2)));
      const float4* tp = reinterpret_cast<const float4*>(tris + ti);
      float4 a = tp[0], b = tp[1], e = tp[2];
      float t, b1, b2;
      bool front;
      if (!intersect_tri(wr, f4xyz(a), f4xyz(b), f4xyz(e), tmin, tmax, t, b1, b2, front)) continue;
      uint32_t gid = __float_as_uint(a.w);
      uint32_t fl = __float_as_uint(b.w) & fmask;
      float aw = e.w;
      if (ANY) {
        if (gid == skip) continue;
        if (ALPHA && !tex_alpha_filter(fl, gid, t, b1, b2, aw)) continue;
        if (!accept_tri(fl, aw, front, false, coin_key, gid)) continue;
        hit_any = true;
        return true;
      } else {
        if (best != HIT_MISS && !(t < tmax || (t == tmax && gid < (best & 0x7fffffffu)))) continue;
        if (ALPHA && !tex_alpha_filter(fl, gid, t, b1, b2, aw)) continue;
        if (!accept_tri(fl, aw, front, cull, coin_key, gid)) continue;
        tmax = t;
        best = gid | (front ? 0u : HIT_BACK);
        bu = b1;
        bv = b2;
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (cnt[c] > 0) tn[c] = __builtin_huge_valf();  // leaves are never pushed
    // Inner children near to far (the nearest is visited first). Occlusion
    // rays too: the nearest blockers end them sooner (-17% any-hit time on
    // the terrain).
    {
      uint32_t ci[4] = {chl[0], chl[1], chl[2], chl[3]};
  // Compare-swap as selects (as a branch, each of the five cost an exec
  // mask save and restore).
#define OXR_CSWAP(i, j)                        \
  {                                            \
    const bool sw = tn[i] < tn[j];             \
    const float ti = tn[i], tj = tn[j];        \
    const uint32_t cii = ci[i], cij = ci[j];   \
    tn[i] = sw ? tj : ti;                      \
    tn[j] = sw ? ti : tj;                      \
    ci[i] = sw ? cij : cii;                    \
    ci[j] = sw ? cii : cij;                    \
  }
      OXR_CSWAP(0, 1) OXR_CSWAP(2, 3) OXR_CSWAP(0, 2) OXR_CSWAP(1, 3) OXR_CSWAP(1, 2)
#undef OXR_CSWAP
      // Sorted far to near, the children to visit are a suffix of ci: the
      // nearest (ci[3]) is visited next straight from a register, and only
      // the others go on the stack. Pushing it and popping it back (the
      // same order) put an LDS store and load between a node's test and
      // the next node's fetch.
      const float lim = tmax * SLAB_EPS;
      const uint32_t k = (tn[0] != __builtin_huge_valf() && tn[0] <= lim ? 1u : 0u) +
                         (tn[1] != __builtin_huge_valf() && tn[1] <= lim ? 1u : 0u) +
                         (tn[2] != __builtin_huge_valf() && tn[2] <= lim ? 1u : 0u);
      st.push_suffix(ci[0], ci[1], ci[2], k);
      if (st.overflow) return true;
      if (tn[3] != __builtin_huge_valf() && tn[3] <= lim) {
        node = ci[3];
        return false;
      }
    }
    return !st.pop(node);
  }
};

// Persistent traversal loop shared by the closest-hit and occlusion
// kernels. A wave refills from the shared counter (one atomic per refill)
// once `refill` of its lanes are idle; the first rays of every wave are
// assigned by wave index, so a launch whose rays fit in one round makes no
// atomic at all. The idle lanes take the next rays of the queue together,
// so rays enter a wave in runs of queue neighbours (coherent node and
// triangle fetches). Waiting for all 64 lanes holds a wave on its slowest
// ray; refilling at a few idle lanes mixes far-apart rays (at 1 idle lane
// 1.4x slower on the material sweep). The engine's defaults, 56 (closest
// hit) and 48 (occlusion), measured 3-5% faster than 64 on the material
// sweep and the terrain; 8-40 were slower.

// nseg > 1: the queue is split into nseg contiguous segments and block b
// works on segment b % nseg. Workgroups are placed on the XCDs round-robin,
// so with nseg = 8 each XCD traces its own run of neighbouring rays and
// keeps their BVH nodes in its own L2. `work` holds one counter per segment;
// gridDim.x is a multiple of nseg.
// segw = nseg | refill << 8: a wave takes new rays once `refill` of its
// lanes are idle (64 when the field is 0).
template <bool ANY, bool ALPHA, bool PATH, class Setup, class Step, class Finish>
OXR_D void trace_loop(uint32_t count, uint32_t* work, uint32_t segw, Stack& st, Setup setup, Step step,
                      Finish finish) {
  const uint32_t lane = lane_id();
  const uint32_t nseg = segw & 0xffu;
  const uint32_t refill = (segw >> 8) ? (segw >> 8) : 64u;
  const uint32_t seg = blockIdx.x % nseg;
  const uint32_t nwaves = (gridDim.x / nseg) * (TRACE_TB / 64);
  const uint32_t wave = (blockIdx.x / nseg) * (TRACE_TB / 64) + threadIdx.x / 64;
  const uint32_t s0 = (uint32_t)((uint64_t)count * seg / nseg);
  const uint32_t s1 = (uint32_t)((uint64_t)count * (seg + 1) / nseg);
  uint32_t* wk = work + (size_t)seg * PAD_COUNTER_WORDS;
  Trav<ANY, ALPHA, PATH> tr;
  tr.asc = nullptr;
  uint32_t ri = s0 + wave * 64u + lane;
  bool active = ri < s1;
  bool exhausted = (uint64_t)s0 + (uint64_t)nwaves * 64u >= s1;
  if (active) setup(tr, ri);
  for (;;) {
    if (active) {
      if (step(tr)) {
        finish(tr, ri);
        active = false;
      }
    }
    uint64_t freem = __ballot(!active);
    uint32_t nfree = (uint32_t)__popcll(freem);
    if (exhausted) {
      if (nfree == 64u) break;
      continue;
    }
    if (nfree >= refill) {
      int leader = __ffsll((unsigned long long)freem) - 1;
      uint32_t base = 0;
      if ((int)lane == leader) base = atomicAdd(wk, nfree) + nwaves * 64u + s0;
      base = __shfl(base, leader);
      if ((uint64_t)base + nfree >= s1) exhausted = true;
      if (!active) {
        uint64_t below = lane == 0 ? 0ull : (freem & ((~0ull) >> (64 - lane)));
        ri = base + (uint32_t)__popcll(below);
        if (ri < s1) {
          setup(tr, ri);
          active = true;
        }
      }
    }
  }
}

// Persistent closest-hit kernel over a ray queue. PATH: rays carry the
// bounce index k in the t_min slot (t_min = 0); else t_min as given.
// MULTI: the launch mixes scenes (server rounds); each ray finds its scene
// through its path's job. Otherwise every ray traces `sc`.
// Occupancy of the traversal kernels: 7 waves per SIMD (72 VGPRs) measured
// 4% faster than the register-limited 6 on the material sweep and 3% on
// the terrain. At 72 VGPRs the path kernels spill only on cold paths (the
// BLEND coin, the stack's global spill, refills); 8 waves (64 VGPRs) put
// reloads back on the leaf test. The textured-alpha variants hold more
// state and keep 4.
#define OXR_TRACE_WAVES(ALPHA) __attribute__((amdgpu_waves_per_eu((ALPHA) ? 4 : 7)))

// COUNT: adds the nodes and triangles the launch fetched to counts[0] and
// counts[32] (two 256-byte lines), one atomic per wave and counter.
template <bool PATH, bool MULTI, bool ALPHA, bool COUNT = false>
__global__ __launch_bounds__(TRACE_TB) OXR_TRACE_WAVES(ALPHA) void k_closest(DevScene sc, const DevJob* __restrict__ jobs,
                                                     const uint32_t* __restrict__ pjob, const float4* __restrict__ rays,
                                                     const uint32_t* __restrict__ qpid,
                                                     const uint4* __restrict__ meta, uint32_t fixed_count,
                                                     const uint32_t* count_ptr, uint32_t* work, uint32_t cull_all,
                                                     uint4* __restrict__ hits, uint32_t* spill, uint32_t spill_cap,
                                                     uint32_t* err, uint32_t nseg, const DevScene* asc,
                                                     unsigned long long* counts) {
  __shared__ uint32_t s_stack[STACK * TRACE_TB];
  Stack st;
  st.lds = s_stack + threadIdx.x;
  st.spill = spill + (size_t)(blockIdx.x * TRACE_TB + threadIdx.x) * spill_cap;
  st.cap = spill_cap;
  const uint32_t count = count_ptr ? *count_ptr : fixed_count;
  uint32_t nnode = 0, ntri = 0;  // COUNT: this lane's fetches
  const Node4Q* ray_nodes = sc.nodes;
  const TriRec* ray_tris = sc.tris;
  auto setup = [&](Trav<false, ALPHA, PATH>& tr, uint32_t i) {
    float4 a = rays[2 * i], b = rays[2 * i + 1];
    uint32_t coin_key = 0;
    float tmin;
    bool cull;
    if (PATH) {
      uint32_t k = __float_as_uint(a.w);
      uint32_t pid = qpid[i];
      bool blend = sc.has_blend;
      if (MULTI) {
        const DevScene& js = jobs[pjob[pid]].sc;
        ray_nodes = js.nodes;
        ray_tris = js.tris;
        blend = js.has_blend;
        if (ALPHA) tr.asc = &js;
      }
      tmin = 0.0f;
      cull = k == 0;
      if (blend) {
        uint4 m = meta[pid];
        coin_key = hash_mix(hash_mix(m.w, m.y), k << 16);
      }
    } else {
      tmin = a.w;
      cull = cull_all != 0;
    }
    if (ALPHA && !MULTI) tr.asc = asc;
    tr.setup(f4xyz(a), f4xyz(b), tmin, b.w, cull, NO_SKIP, coin_key, st);
  };
  auto step = [&](Trav<false, ALPHA, PATH>& tr) {
    if (COUNT) ++nnode;
    return MULTI ? tr.step(st, ray_nodes, ray_tris, ntri) : tr.template step<COUNT>(st, sc.nodes, sc.tris, ntri);
  };
  auto finish = [&](Trav<false, ALPHA, PATH>& tr, uint32_t i) {
    uint32_t prim = tr.best;
    if (st.overflow) {
      atomicOr(err, ERR_STACK);
      prim = HIT_MISS;
    }
    hits[i] = make_uint4(prim == HIT_MISS ? 0u : __float_as_uint(tr.tmax), prim, __float_as_uint(tr.bu),
                         __float_as_uint(tr.bv));
  };
  trace_loop<false, ALPHA, PATH>(count, work, nseg, st, setup, step, finish);
  if (COUNT) {  // every lane of the wave leaves trace_loop together
    for (int o = 32; o > 0; o >>= 1) {
      nnode += __shfl_xor(nnode, o);
      ntri += __shfl_xor(ntri, o);
    }
    if (lane_id() == 0u) {
      atomicAdd(counts, (unsigned long long)nnode);
      atomicAdd(counts + 32, (unsigned long long)ntri);
    }
  }
}

// Persistent occlusion kernel. PATH: record = (o, skip gid), (d, t_max),
// sidx[i] = path * nslots + slot; writes vis[slot] = 2 when unoccluded.
// Else: record = oxr_ray, writes occ[i] (1 = occluded). MULTI as k_closest.
// COUNT: as k_closest (nodes fetched, triangles tested up to the first
// accepted one).
template <bool PATH, bool MULTI, bool ALPHA, bool COUNT = false>
__global__ __launch_bounds__(TRACE_TB) OXR_TRACE_WAVES(ALPHA) void k_anyhit(DevScene sc, const DevJob* __restrict__ jobs,
                                                    const uint32_t* __restrict__ pjob, const float4* __restrict__ rays,
                                                    const uint32_t* __restrict__ sidx, const uint4* __restrict__ meta,
                                                    uint32_t nslots, uint32_t fixed_count, const uint32_t* count_ptr,
                                                    uint32_t* work, uint8_t* __restrict__ out, uint32_t* spill,
                                                    uint32_t spill_cap, uint32_t* err, uint32_t nseg,
                                                    const DevScene* asc, unsigned long long* counts) {
  __shared__ uint32_t s_stack[STACK * TRACE_TB];
  Stack st;
  st.lds = s_stack + threadIdx.x;
  st.spill = spill + (size_t)(blockIdx.x * TRACE_TB + threadIdx.x) * spill_cap;
  st.cap = spill_cap;
  const uint32_t count = count_ptr ? *count_ptr : fixed_count;
  uint32_t nnode = 0, ntri = 0;  // COUNT: this lane's fetches
  const Node4Q* ray_nodes = sc.nodes;
  const TriRec* ray_tris = sc.tris;
  auto setup = [&](Trav<true, ALPHA, PATH>& tr, uint32_t i) {
    float4 a = rays[2 * i], b = rays[2 * i + 1];
    float tmin = 0.0f;
    uint32_t skip = NO_SKIP, coin_key = 0;
    if (PATH) {
      skip = __float_as_uint(a.w);
      uint32_t slot = sidx[i];
      uint32_t pid = slot / nslots, j = slot - pid * nslots;
      bool blend = sc.has_blend;
      if (MULTI) {
        const DevScene& js = jobs[pjob[pid]].sc;
        ray_nodes = js.nodes;
        ray_tris = js.tris;
        blend = js.has_blend;
        if (ALPHA) tr.asc = &js;
      }
      if (blend) {
        uint4 m = meta[pid];
        // The shading kernel has already advanced the bounce of a path that
        // continues (it sets PF_DONE on the others): the shadow rays belong
        // to the vertex before.
        uint32_t k = (m.z & 0xffffu);
        if (!((m.z >> 24) & PF_DONE)) k -= 1u;
        // Punctual slot j is the j-th light kept at upload; the coin uses
        // the light's index in the scene's full list.
        const DevLight* lts = MULTI ? jobs[pjob[pid]].sc.lights : sc.lights;
        const uint32_t ray_id = (k << 16) | (j == nslots - 1 ? 0xffffu : (0x8000u | (lts[j].index & 0x7fffu)));
        coin_key = hash_mix(hash_mix(m.w, m.y), ray_id);
      }
    } else {
      tmin = a.w;
    }
    if (ALPHA && !MULTI) tr.asc = asc;
    tr.setup(f4xyz(a), f4xyz(b), tmin, b.w, false, skip, coin_key, st);
  };
  auto step = [&](Trav<true, ALPHA, PATH>& tr) {
    if (COUNT) ++nnode;
    return MULTI ? tr.step(st, ray_nodes, ray_tris, ntri) : tr.template step<COUNT>(st, sc.nodes, sc.tris, ntri);
  };
  auto finish = [&](Trav<true, ALPHA, PATH>& tr, uint32_t i) {
    bool occ = tr.hit_any;
    if (st.overflow) {
      atomicOr(err, ERR_STACK);
      occ = true;
    }
    if (PATH) {
      if (!occ) out[sidx[i]] = 2;
    } else {
      out[i] = occ ? 1 : 0;
    }
  };
  trace_loop<true, ALPHA, PATH>(count, work, nseg, st, setup, step, finish);
  if (COUNT) {  // every lane of the wave leaves trace_loop together
    for (int o = 32; o > 0; o >>= 1) {
      nnode += __shfl_xor(nnode, o);
      ntri += __shfl_xor(ntri, o);
    }
    if (lane_id() == 0u) {
      atomicAdd(counts, (unsigned long long)nnode);
      atomicAdd(counts + 32, (unsigned long long)ntri);
    }
  }
}

// ---------------------------------------------------------------------
// Camera rays (Sampler::new + pattern 0 + Camera::primary_ray,
// src/pathtrace.rs:2005-2012, src/camera.rs:489-514).
// ---------------------------------------------------------------------
// Pixel offset of rank r in a band of `rows` whole rows of width w,
// visited in TxT tiles (row-major tiles, row-major inside a tile; narrower
// tiles at the right edge, shorter ones at the bottom): a bijection of
// [0, rows * w).
template <uint32_t T>
OXR_D uint32_t tile_order(uint32_t r, uint32_t w, uint32_t rows) {
  const uint32_t tr = r / (T * w);
  const uint32_t ir = r - tr * T * w;
  const uint32_t h = min(T, rows - T * tr);
  const uint32_t full = w / T;
  const uint32_t c = ir / (T * h);
  uint32_t x, y;
  if (c < full) {
    const uint32_t j = ir - c * T * h;
    y = T * tr + j / T;
    x = T * c + j % T;
  } else {
    const uint32_t tw = w - T * full, j = ir - full * T * h;
    y = T * tr + j / tw;
    x = T * full + j % tw;
  }
  return y * w + x;
}

// Segment of path p (segments are in path order).
OXR_D uint32_t seg_of(const SegTable& segs, uint32_t p) {
  uint32_t si = 0;
  while (si + 1 < segs.n && segs.s[si + 1].path0 <= p) ++si;
  return si;
}

// Path p of a segment is (sample s, pixel offset off). Where the segment
// covers whole rows, consecutive paths walk 4x4-pixel tiles with 4 samples
// of a pixel side by side (64 paths: 16 pixels x 4 samples) when S is a
// multiple of 4, else 8x8-pixel tiles of one sample; so a wave's 64 primary
// rays start close together, and, through the queue order, so do their
// later bounces. Otherwise paths go sample-major in pixel order. Path state
// is indexed by p; k_finalize writes each result at s * npx + off, where
// k_reduce reads it.
OXR_D void path_sample(const Seg& sg, uint32_t width, uint32_t lq, uint32_t& s, uint32_t& off) {
  if (width > 0u && sg.px0 % width == 0u && sg.npx % width == 0u) {
    const uint32_t rows = sg.npx / width;
    if (sg.S % 4u == 0u) {
      const uint32_t g = lq / (4u * sg.npx), rem = lq - g * 4u * sg.npx;
      s = 4u * g + rem % 4u;
      off = tile_order<4>(rem / 4u, width, rows);
    } else {
      s = lq / sg.npx;
      off = tile_order<8>(lq - s * sg.npx, width, rows);
    }
  } else {
    s = lq / sg.npx;
    off = lq - s * sg.npx;
  }
}

// Throughput, radiance, previous position and medium of a camera path are
// not written here: the bounce-0 shading launch starts every path from
// beta = 1, zero radiance and no medium without reading them, and writes
// all of them at its end.
__global__ void k_camera(const DevJob* __restrict__ jobs, SegTable segs, uint32_t P, float4* __restrict__ qray,
                         uint32_t* __restrict__ qpid, uint4* __restrict__ meta, uint32_t* __restrict__ pjob, Ctrl* ctrl) {
  const uint32_t pid = blockIdx.x * blockDim.x + threadIdx.x;
  if (pid == 0) ctrl->qcount[0].v = P;
  if (pid >= P) return;
  const Seg sg = segs.s[seg_of(segs, pid)];
  const DevFrame& fr = jobs[sg.job].fr;
  uint32_t s, off;
  path_sample(sg, fr.win_w, pid - sg.path0, s, off);
  uint32_t px = sg.px0 + off;
  uint32_t x = fr.win_x0 + px % fr.win_w, y = fr.win_y0 + px / fr.win_w;
  uint32_t index = sg.first + s;
  uint32_t ps = pixel_seed_of(fr.seed, x, y);
  F4 p0 = pattern(ps, index, 0);
  float fx = (float)x + p0.v[0], fy = (float)y + p0.v[1];
  float ndc_x = (fx / fmaxr((float)fr.width, 1.0f)) * 2.0f - 1.0f;
  float ndc_y = 1.0f - (fy / fmaxr((float)fr.height, 1.0f)) * 2.0f;
  V3 side = ld3(fr.side), up = ld3(fr.up), fwd = ld3(fr.forward), eye = ld3(fr.eye);
  V3 o, d;
  if (fr.projection == 1u) {
    o = add(eye, add(scale(side, ndc_x * fr.half_w), scale(up, ndc_y * fr.half_h)));
    d = fwd;
  } else {
    d = normalise(add(add(scale(side, ndc_x * fr.half_w), scale(up, ndc_y * fr.half_h)), fwd));
    o = eye;
  }
  qray[2 * pid] = make_float4(o.x, o.y, o.z, __uint_as_float(0u));
  qray[2 * pid + 1] = make_float4(d.x, d.y, d.z, __builtin_huge_valf());
  qpid[pid] = pid;
  meta[pid] = make_uint4(x | (y << 16), index, 0xffu << 16, ps);
  pjob[pid] = sg.job;
}

__global__ void k_prep(Ctrl* ctrl, uint32_t cur, uint32_t count_shadow_prev) {
  if (threadIdx.x == 0) {
    ctrl->rays_closest += ctrl->qcount[cur].v;
    if (count_shadow_prev) ctrl->rays_shadow += ctrl->scount.v;
    ctrl->qcount[cur ^ 1u].v = 0;
    ctrl->scount.v = 0;
  }
  if (threadIdx.x < 16) ctrl->work[threadIdx.x].v = 0;
}

// Adds the visible NEE contributions of the previous vertex in light
// order (punctual lights, then the area sample), as the reference adds
// them to `radiance` during that vertex.
OXR_D void add_pending(uint32_t pid, uint32_t nslots, const float4* __restrict__ nee, uint8_t* __restrict__ vis,
                       V3& radiance) {
  for (uint32_t j = 0; j < nslots; ++j) {
    size_t s = (size_t)pid * nslots + j;
    uint8_t v = vis[s];
    if (v == 2) radiance = add(radiance, f4xyz(nee[s]));
    if (v) vis[s] = 0;
  }
}

OXR_D V3 clampc(V3 c, float cl) {
  float m = max3(c);
  if (cl > 0.0f && m > cl) return scale(c, cl / m);
  return c;
}

// Light-sampling pdf of emissive triangle g (Integrator::tri_pdf,
// src/pathtrace.rs:1972-1991), given its corners and its solid angle from
// p (the area sample computes both before it picks a direction; tri_pdf
// below computes them).
OXR_D float tri_pdf_of(const DevScene& sc, float p_env, uint32_t li, V3 a, V3 b, V3 c, float omega, V3 p, V3 q) {
  float sel = (1.0f - p_env) * sc.l_pmf[li];
  if (omega >= 1.0e-4f) return sel / omega;
  V3 cr = cross(sub(b, a), sub(c, a));
  float len = sqrtf(dot(cr, cr));
  V3 to = sub(q, p);
  float d2 = dot(to, to);
  if (len <= 0.0f || d2 <= 0.0f) return 0.0f;
  float cos = fabsf(dot(cr, to) / (len * sqrtf(d2)));
  if (cos <= 0.0f) return 0.0f;
  return sel * d2 / (0.5f * len * cos);
}
OXR_D float tri_pdf(const DevScene& sc, float p_env, uint32_t li, uint32_t g, V3 p, V3 q) {
  const float* tv = sc.pos + 9ull * g;
  V3 a = ld3(tv), b = ld3(tv + 3), c = ld3(tv + 6);
  return tri_pdf_of(sc, p_env, li, a, b, c, triangle_solid_angle(p, a, b, c), p, q);
}

// ---------------------------------------------------------------------
// Environment map (EnvironmentMap::{texel_of, radiance, pdf, sample},
// src/pathtrace.rs:1677-1736).
// ---------------------------------------------------------------------
OXR_D void env_texel(const DevEnv& e, V3 d, uint32_t& x, uint32_t& y, float& st) {
  float u = 0.5f + atan2f(d.x, -d.z) / (2.0f * PI_F);
  float th = acosf(clampr(d.y, -1.0f, 1.0f));
  float v = th / PI_F;
  float fx = u * (float)e.w, fy = v * (float)e.h;
  x = fx > 0.0f ? (uint32_t)fx : 0u;
  y = fy > 0.0f ? (uint32_t)fy : 0u;
  x = x < e.w - 1u ? x : e.w - 1u;
  y = y < e.h - 1u ? y : e.h - 1u;
  st = sinf(th);
}
OXR_D V3 env_radiance(const DevEnv& e, V3 d) {
  uint32_t x, y;
  float st;
  env_texel(e, d, x, y, st);
  if (e.uniform) return v3(0.0f, 0.0f, 0.0f);
  float4 c = e.img[(size_t)y * e.w + x];
  if (!(isfinite(c.x) && isfinite(c.y) && isfinite(c.z))) return v3(0.0f, 0.0f, 0.0f);
  return scale(v3(fmaxr(c.x, 0.0f), fmaxr(c.y, 0.0f), fmaxr(c.z, 0.0f)), e.intensity);
}
OXR_D float env_pdf(const DevEnv& e, V3 d) {
  uint32_t x, y;
  float st;
  env_texel(e, d, x, y, st);
  if (st <= 1.0e-6f) return 0.0f;
  return e.mass[(size_t)y * e.w + x] * (float)(e.w * e.h) / (2.0f * PI_F * PI_F * st);
}
// partition_point(c <= u) over n entries.
OXR_D uint32_t cdf_point(const float* c, uint32_t n, float u) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (c[mid] <= u)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}
OXR_D bool env_sample(const DevEnv& e, float u0, float u1, V3& d, V3& le, float& pdf) {
  uint32_t y = cdf_point(e.marginal, e.h + 1u, u0);
  y = (y < 1u ? 1u : (y > e.h ? e.h : y)) - 1u;
  const float* row = e.conditional + (size_t)y * (e.w + 1u);
  uint32_t x = cdf_point(row, e.w + 1u, u1);
  x = (x < 1u ? 1u : (x > e.w ? e.w : x)) - 1u;
  float my = (u0 - e.marginal[y]) / fmaxr(e.marginal[y + 1] - e.marginal[y], 1e-12f);
  float mx = (u1 - row[x]) / fmaxr(row[x + 1] - row[x], 1e-12f);
  float u = ((float)x + clampr(mx, 0.0f, 0.99999f)) / (float)e.w;
  float v = ((float)y + clampr(my, 0.0f, 0.99999f)) / (float)e.h;
  float phi = 2.0f * PI_F * (u - 0.5f);
  float th = PI_F * v;
  float st = sinf(th), ct = cosf(th);
  d = v3(st * sinf(phi), ct, -st * cosf(phi));
  pdf = env_pdf(e, d);
  if (!(pdf > 0.0f) || !isfinite(pdf)) return false;
  le = env_radiance(e, d);
  return true;
}

OXR_D void load_mat(const DevMaterial& m, V3 sn, MatSample& ms) {
  ms.base = v3(m.base[0], m.base[1], m.base[2]);
  ms.metallic = m.metallic;
  ms.roughness = m.roughness;
  ms.normal = sn;
  ms.ior = m.ior;
  ms.specular = m.specular;
  ms.specular_color = ld3(m.spec_color);
  ms.transmission = m.transmission;
  ms.thickness = m.thickness;
  ms.clearcoat = m.clearcoat;
  ms.clearcoat_roughness = m.cc_rough;
  ms.clearcoat_normal = sn;
  ms.sheen_color = ld3(m.sheen);
  ms.sheen_roughness = m.sheen_rough;
  ms.anisotropy = 0.0f;
  ms.anisotropy_direction = v3(0.0f, 0.0f, 0.0f);
  ms.iridescence = m.iridescence;
  ms.iridescence_ior = m.irid_ior;
  ms.iridescence_thickness = m.irid_thick;
  ms.diffuse_transmission = m.dt;
  ms.diffuse_transmission_color = ld3(m.dt_color);
}

// The textured reads of TraceScene::material_oriented (src/trace.rs:
// 717-895) on top of the factors load_mat put in `ms`. snf: the shading
// normal, already reversed on a double-sided back face (`flip`); the
// mapped normals are reversed back afterwards (src/pathtrace.rs:2093-2101).
OXR_D void tex_apply(const DevTextures& T, uint2 run, uint32_t g, uint32_t attr, V3 bary, const float* pos9,
                     const TexCone& cone, V3 snf, bool flip, const DevMaterial& M, MatSample& ms) {
  V3 normal = snf, ccn = snf;
  float irid_thick = ms.iridescence_thickness;
  for (uint32_t i = 0; i < run.y; ++i) {
    const DevTexBind b = T.binds[run.x + i];
    F4 t;
    if (!tex_tap(T, b, g, attr, bary, pos9, cone, t)) continue;
    switch (b.slot) {
      case TS_BASE: ms.base = v3(ms.base.x * t.v[0], ms.base.y * t.v[1], ms.base.z * t.v[2]); break;
      case TS_MR:
        ms.roughness *= t.v[1];
        ms.metallic *= t.v[2];
        break;
      case TS_NORMAL: normal = tex_perturb(T, g, attr, bary, snf, t, b.aux); break;
      case TS_SPEC: ms.specular *= t.v[3]; break;
      case TS_SPEC_COLOR: ms.specular_color = mul(ms.specular_color, v3(t.v[0], t.v[1], t.v[2])); break;
      case TS_TRANSMISSION: ms.transmission *= t.v[0]; break;
      case TS_THICKNESS: ms.thickness *= t.v[1]; break;
      case TS_CC: ms.clearcoat *= t.v[0]; break;
      case TS_CC_ROUGH: ms.clearcoat_roughness *= t.v[1]; break;
      case TS_CC_NORMAL: ccn = tex_perturb(T, g, attr, bary, snf, t, b.aux); break;
      case TS_SHEEN_COLOR: ms.sheen_color = mul(ms.sheen_color, v3(t.v[0], t.v[1], t.v[2])); break;
      case TS_SHEEN_ROUGH: ms.sheen_roughness *= t.v[3]; break;
      case TS_IRID: ms.iridescence *= t.v[0]; break;
      case TS_IRID_THICK: {
        float lo = b.aux, hi = M.irid_thick;
        irid_thick = fmaxr(lo + (hi - lo) * clampr(t.v[1], 0.0f, 1.0f), 0.0f);
        break;
      }
      case TS_DT: ms.diffuse_transmission *= t.v[3]; break;
      case TS_DT_COLOR:
        ms.diffuse_transmission_color = mul(ms.diffuse_transmission_color, v3(t.v[0], t.v[1], t.v[2]));
        break;
      default: break;  // emissive: Integrator::emission; anisotropy: below
    }
  }
  if (attr & TA_COLORS) {
    F4 c = tex_interp4(T.col + 12ull * g, bary);
    ms.base = v3(ms.base.x * c.v[0], ms.base.y * c.v[1], ms.base.z * c.v[2]);
  }
  ms.metallic = clampr(ms.metallic, 0.0f, 1.0f);
  ms.roughness = clampr(ms.roughness, 0.0f, 1.0f);
  ms.iridescence_thickness = irid_thick;
  // KHR_materials_anisotropy (src/trace.rs:835-858).
  if (M.aniso > 0.0f) {
    float a = M.aniso;
    float d0 = 1.0f, d1 = 0.0f;
    const DevTexBind* ab = tex_find(T, run, TS_ANISO);
    F4 t;
    if (ab && tex_tap(T, *ab, g, attr, bary, pos9, cone, t)) {
      d0 = t.v[0] * 2.0f - 1.0f;
      d1 = t.v[1] * 2.0f - 1.0f;
      a *= clampr(t.v[2], 0.0f, 1.0f);
    }
    float sr = sinf(M.aniso_rot), cr = cosf(M.aniso_rot);
    float r0 = cr * d0 - sr * d1, r1 = sr * d0 + cr * d1;
    V3 dir = v3(0.0f, 0.0f, 0.0f), tt, bt;
    if (tex_tangent_frame(T, g, attr, bary, snf, tt, bt)) {
      V3 d = normalise(v3(tt.x * r0 + bt.x * r1, tt.y * r0 + bt.y * r1, tt.z * r0 + bt.z * r1));
      if (dot(d, d) > 0.5f) dir = d;
    }
    ms.anisotropy = dot(dir, dir) < 0.5f ? 0.0f : a;
    ms.anisotropy_direction = dir;
  }
  ms.normal = flip ? neg(normal) : normal;
  ms.clearcoat_normal = flip ? neg(ccn) : ccn;
}

OXR_D V3 spawn(V3 p, V3 ng, V3 dir) { return dot(dir, ng) >= 0.0f ? offset_ray_origin(p, ng) : offset_ray_origin(p, neg(ng)); }

// ---------------------------------------------------------------------
// Shading: one path vertex of Integrator::sample (src/pathtrace.rs:
// 2027-2322), with shadow rays deferred to k_anyhit.
// ---------------------------------------------------------------------
// MULTI: the batch mixes jobs (each path's job through pjob); else every
// path belongs to job 0, whose scene and frame the wave reads with uniform
// (scalar) loads instead of a per-lane dependent lookup.
// AREA = false: no scene of the batch has an emissive triangle in its light
// table and no job has an environment map, so the area-light sample, the
// light-table MIS weight of an emitter hit (it has none: pl = 0, w = 1) and
// the environment lookups are never taken and are left out.
// EXT = false: no material of the batch has an extension lobe (clearcoat,
// sheen, iridescence, transmission, diffuse transmission, anisotropy), so
// no path ever enters a medium: the medium state stays zero and is neither
// read nor written, and the BSDF drops those lobes (BsdfT<0>). EXT = 1:
// no material of the batch has iridescence (BsdfT<1>); EXT = 2: all lobes.
// PUNCT = false: the batch has no punctual light (nslots == 1), so the
// punctual NEE loop, whose inlined BSDF evaluation held the most registers
// of the extension + area variant, is left out.
// With at most STAGE_SLOTS shadow slots per path (punctual lights + the
// area sample) a block stages its round's shadow rays and next rays in LDS
// and appends each queue with one atomic per round (block_flush); one
// atomic per wave and light made the counter the kernel's bottleneck.
constexpr uint32_t STAGE_SLOTS = 4;

template <int WAVES, bool GUIDE, bool TEX, bool MULTI, int EXT, bool AREA, bool PUNCT = true>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_shade(const DevJob* __restrict__ jobs,
                                             const uint32_t* __restrict__ pjob, uint32_t nslots, const float4* __restrict__ qray,
                                             const uint32_t* __restrict__ qpid, const uint4* __restrict__ hits,
                                             float4* __restrict__ nray, uint32_t* __restrict__ npid,
                                             float4* __restrict__ sray, uint32_t* __restrict__ sidx,
                                             float4* __restrict__ beta_s, float4* __restrict__ rad_s,
                                             float4* __restrict__ prev_s, float4* __restrict__ sigma_s,
                                             uint4* __restrict__ meta_s, float4* __restrict__ nee,
                                             uint8_t* __restrict__ vis, Ctrl* ctrl, uint32_t cur,
                                             DevGuideUse gu, uint32_t kb) {
  __shared__ float4 s_sa[STAGE_SLOTS * 256], s_sb[STAGE_SLOTS * 256], s_na[256], s_nb[256];
  __shared__ uint32_t s_si[STAGE_SLOTS * 256], s_np[256], s_cnt[8], s_base[8];
  // Uniform over the grid: every thread reaches each block_flush.
  const bool stage = nslots <= STAGE_SLOTS;
  const uint32_t wbase = (threadIdx.x / 64u) * STAGE_SLOTS * 64u;  // this wave's shadow staging
  const uint32_t nbase = (threadIdx.x / 64u) * 64u;                // this wave's next-ray staging
  const uint32_t count = ctrl->qcount[cur].v;
  const uint32_t rec_mask = GUIDE ? (1u << gu.shift) - 1u : 0u;
  uint32_t* ncount = &ctrl->qcount[cur ^ 1u].v;
  uint32_t* scount = &ctrl->scount.v;
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
  // Uniform trip count so every lane reaches the wave-level appends.
  const uint32_t rounds = (count + stride - 1) / stride;
  // Software pipelining: the next round's queue entry (path id, hit, ray
  // direction) is in flight while this round shades, so its dependent
  // loads (path state, triangle, material) start one memory latency
  // earlier.
  uint32_t pid_n = i0 < count ? qpid[i0] : 0u;
  uint4 h_n = i0 < count ? hits[i0] : make_uint4(0, HIT_MISS, 0, 0);
  float4 rd_n = i0 < count ? qray[2 * i0 + 1] : make_float4(0, 0, 1, 0);
  // A thread past the queue's end shades nothing but still reads a job's
  // scene (the material loads below are unconditional). It takes the job
  // of the queue's first path, which is rendering, not a job-table slot
  // that may hold a finished job whose scene the server has freed.
  const uint32_t pid_any = (MULTI && count > 0u) ? qpid[0] : 0u;
  for (uint32_t r = 0; r < rounds; ++r) {
    uint32_t i = i0 + r * stride;
    bool valid = i < count;
    uint32_t pid = valid ? pid_n : 0u;
    const uint4 h_cur = h_n;
    const float4 rd_cur = rd_n;
    {
      uint32_t in = i + stride;
      if (r + 1 < rounds && in < count) {
        pid_n = qpid[in];
        h_n = hits[in];
        rd_n = qray[2 * in + 1];
      }
    }
    bool alive = valid;
    uint32_t staged = 0, nstaged = 0;  // entries this wave staged this round (wave-uniform)
    const DevJob& job = jobs[MULTI ? pjob[valid ? pid : pid_any] : 0u];
    const DevScene& sc = job.sc;
    const DevFrame& fr = job.fr;
    uint4 meta = valid ? meta_s[pid] : make_uint4(0, 0, 0, 0);
    uint32_t k = meta.z & 0xffffu;
    uint32_t hero = (meta.z >> 16) & 0xffu;
    uint32_t flags = meta.z >> 24;
    uint32_t ps = meta.w, index = meta.y;
    // kb: the launch's bounce (every queue entry of a launch has it). At
    // bounce 0 the path starts from beta = 1 and nothing else, unread.
    const bool later = valid && kb > 0u;
    float4 b4 = later ? beta_s[pid] : (valid ? make_float4(1.0f, 1.0f, 1.0f, 0.0f) : make_float4(0, 0, 0, 0));
    V3 beta = f4xyz(b4);
    float prev_pdf = b4.w;
    V3 radiance = later ? f4xyz(rad_s[pid]) : v3(0, 0, 0);
    float4 sg = (EXT && later) ? sigma_s[pid] : make_float4(0, 0, 0, 0);
    float4 pv4 = later ? prev_s[pid] : make_float4(0, 0, 0, 0);
    V3 prev_pos = f4xyz(pv4);
    // Roulette throughput scale of guided paths: the product of
    // p_mixture / p_bsdf over the guided vertices so far (1 unguided).
    float rr_scale = (GUIDE && k > 0) ? pv4.w : 1.0f;
    if (valid && (flags & PF_PENDING)) {
      add_pending(pid, nslots, nee, vis, radiance);
      flags &= ~PF_PENDING;
    }
    // Guide training record of the previous vertex: the path radiance now
    // holds everything that vertex added itself (emission, NEE).
    if (GUIDE && gu.rec && valid && k >= 1 && k - 1 < (uint32_t)G_KREC && (pid & rec_mask) == 0u)
      gu.rec[((size_t)(pid >> gu.shift) * G_KREC + (k - 1)) * 4 + 3] =
          make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
    uint4 h = valid ? h_cur : make_uint4(0, HIT_MISS, 0, 0);
    float4 rd = valid ? rd_cur : make_float4(0, 0, 1, 0);
    V3 d = f4xyz(rd);
    bool miss = h.y == HIT_MISS;
    float t = __uint_as_float(h.x);
    // Beer-Lambert inside a medium; an escaping ray ends the path.
    if (EXT && alive && sg.w != 0.0f) {
      if (miss) {
        alive = false;
      } else {
        beta.x *= expf(-sg.x * t);
        beta.y *= expf(-sg.y * t);
        beta.z *= expf(-sg.z * t);
      }
    }
    if (alive && miss) {
      if (k == 0) {
        flags |= PF_UNCOVERED;
      } else {
        // The environment map when installed, else the constant sky
        // (src/pathtrace.rs:2038-2050).
        V3 le = v3(fr.ambient, fr.ambient, fr.ambient);
        float w = 1.0f;
        if (AREA && fr.env.img) {
          le = env_radiance(fr.env, d);
          w = fr.strategy == 0u ? power_heuristic(prev_pdf, fr.p_env * env_pdf(fr.env, d))
                                : (fr.strategy == 1u ? 0.0f : 1.0f);
        }
        if (w > 0.0f) radiance = add(radiance, clampc(scale(mul(beta, le), w), fr.clamp));
      }
      alive = false;
    }
    // Surface + emission.
    uint32_t g = h.y & 0x7fffffffu;
    bool front = !(h.y & HIT_BACK);
    V3 p = v3(0, 0, 0), ngeo = v3(0, 0, 1), sn = v3(0, 0, 1);
    V3 bary = v3(1.0f, 0.0f, 0.0f);
    uint32_t mid = 0;
    if (alive) {
      float u = __uint_as_float(h.z), v = __uint_as_float(h.w);
      bary = v3(1.0f - u - v, u, v);
      const float* tv = sc.pos + 9ull * g;
      V3 p0 = ld3(tv), p1 = ld3(tv + 3), p2 = ld3(tv + 6);
      p = v3(p0.x * bary.x + p1.x * bary.y + p2.x * bary.z, p0.y * bary.x + p1.y * bary.y + p2.y * bary.z,
             p0.z * bary.x + p1.z * bary.y + p2.z * bary.z);
      V3 cr = cross(sub(p1, p0), sub(p2, p0));
      float len = sqrtf(dot(cr, cr));
      ngeo = len > 0.0f ? v3(cr.x / len, cr.y / len, cr.z / len) : v3(0.0f, 0.0f, 1.0f);
      uint32_t info = sc.tri_info[g];
      mid = info & TI_MAT_MASK;
      sn = ngeo;
      if (info & TI_HAS_NORMALS) {
        const float* nv = sc.nrm + 9ull * g;
        V3 n0 = ld3(nv), n1 = ld3(nv + 3), n2 = ld3(nv + 6);
        V3 n = normalise(v3(n0.x * bary.x + n1.x * bary.y + n2.x * bary.z, n0.y * bary.x + n1.y * bary.y + n2.y * bary.z,
                            n0.z * bary.x + n1.z * bary.y + n2.z * bary.z));
        sn = dot(n, n) > 0.5f ? n : ngeo;
      }
    }
    const DevMaterial* M = sc.mats + mid;
    if (alive) {
      V3 em = ld3(M->emissive);
      if (TEX && !is_black(em)) em = mul(em, tex_emission_factor(sc.tex, mid, g, bary));
      bool ds = (M->flags & MAT_DS) != 0;
      bool unlit = (M->flags & MAT_UNLIT) != 0;
      if ((front || ds) && !unlit && !is_black(em)) {
        if (k == 0) {
          radiance = add(radiance, mul(beta, em));
        } else {
          uint32_t li = AREA ? sc.l_lookup[g] : 0xffffffffu;
          float pl = li == 0xffffffffu ? 0.0f : tri_pdf(sc, fr.p_env, li, g, prev_pos, p);
          float w;
          if (pl <= 0.0f)
            w = 1.0f;
          else if (fr.strategy == 0u)
            w = power_heuristic(prev_pdf, pl);
          else if (fr.strategy == 1u)
            w = 0.0f;
          else
            w = 1.0f;
          if (w > 0.0f) radiance = add(radiance, clampc(scale(mul(beta, em), w), fr.clamp));
        }
      }
      if (unlit) {
        V3 c = v3(M->base[0], M->base[1], M->base[2]);
        V3 contrib = mul(beta, c);
        radiance = add(radiance, k == 0 ? contrib : clampc(contrib, fr.clamp));
        alive = false;
      }
    }
    if (alive && k >= fr.max_bounces) alive = false;
    // The area set's NEE sample (src/pathtrace.rs:2173-2279) is drawn here,
    // before the material and the BSDF exist: it needs only the hit point,
    // and its f64 spherical-triangle sampling then does not hold registers
    // beside the BSDF's lobes (which spilled 400 bytes a lane in the
    // extension + area variant). The BSDF evaluation, MIS weight and shadow
    // ray follow once the BSDF is built. Every operation is the one the
    // combined order runs, on the same operands, so results are unchanged.
    bool a_have = false;
    V3 a_l = v3(0, 0, 0), a_le = v3(0, 0, 0);
    float a_pl = 0.0f, a_tmax = 0.0f;
    uint32_t a_skip = NO_SKIP;
    if (AREA && alive && fr.strategy != 2u && (fr.env.img != nullptr || sc.num_emissive > 0)) {
      const float pe = fr.p_env;
      const F4 pl_u = pattern(ps, index, 2 + 2 * k);
      if (pl_u.v[2] < pe) {
        float epdf;
        if (env_sample(fr.env, pl_u.v[0], pl_u.v[1], a_l, a_le, epdf)) {
          a_pl = pe * epdf;
          a_tmax = __builtin_huge_valf();
          a_have = !is_black(a_le) && a_pl > 0.0f;
        }
      } else if (sc.num_emissive > 0) {
        float us = pe > 0.0f ? clampr((pl_u.v[2] - pe) / (1.0f - pe), 0.0f, 0.999999f) : pl_u.v[2];
        // EmissiveLights::pick: partition_point(c <= u).min(len - 1).
        uint32_t lo = 0, hi = sc.num_emissive;
        while (lo < hi) {
          uint32_t mid2 = (lo + hi) >> 1;
          if (sc.l_cdf[mid2] <= us)
            lo = mid2 + 1;
          else
            hi = mid2;
        }
        uint32_t li = lo < sc.num_emissive - 1 ? lo : sc.num_emissive - 1;
        uint32_t lg = sc.l_tris[li];
        const float* tv = sc.pos + 9ull * lg;
        V3 a = ld3(tv), b = ld3(tv + 3), c = ld3(tv + 6);
        bool have = false;
        V3 q;
        // One solid angle serves the choice of sampler and the pdf below
        // (tri_pdf recomputes the same value from the same operands).
        const float omega = triangle_solid_angle(p, a, b, c);
        if (omega >= 1.0e-4f) {
          V3 ls;
          if (sample_spherical_triangle(p, a, b, c, pl_u.v[0], pl_u.v[1], ls)) {
            V3 n = cross(sub(b, a), sub(c, a));
            float dn = dot(ls, n);
            if (dn != 0.0f) {
              float tt = dot(sub(a, p), n) / dn;
              if (tt > 0.0f && isfinite(tt)) {
                q = add(p, scale(ls, tt));
                have = true;
              }
            }
          }
        } else {
          float su = sqrtf(pl_u.v[0]);
          float w0 = 1.0f - su, w1 = su * (1.0f - pl_u.v[1]), w2 = su * pl_u.v[1];
          q = v3(a.x * w0 + b.x * w1 + c.x * w2, a.y * w0 + b.y * w1 + c.y * w2, a.z * w0 + b.z * w1 + c.z * w2);
          have = true;
        }
        V3 lbary;
        if (have && barycentric_of(q, a, b, c, lbary)) {
          V3 to = sub(q, p);
          float dist2 = dot(to, to);
          V3 cr = cross(sub(b, a), sub(c, a));
          float len = sqrtf(dot(cr, cr));
          V3 lng = len > 0.0f ? v3(cr.x / len, cr.y / len, cr.z / len) : v3(0.0f, 0.0f, 1.0f);
          const uint32_t lmid = sc.tri_info[lg] & TI_MAT_MASK;
          const DevMaterial& LM = sc.mats[lmid];
          if (dist2 > 1.0e-12f) {
            float dist = sqrtf(dist2);
            a_l = scale(to, 1.0f / dist);
            float cos_l = -dot(lng, a_l);
            a_pl = tri_pdf_of(sc, fr.p_env, li, a, b, c, omega, p, q);
            bool lds = (LM.flags & MAT_DS) != 0;
            if ((cos_l > 0.0f || (lds && cos_l < 0.0f)) && a_pl > 0.0f && isfinite(a_pl)) {
              a_le = (LM.flags & MAT_UNLIT) ? v3(0, 0, 0) : ld3(LM.emissive);
              if (TEX && !is_black(a_le)) a_le = mul(a_le, tex_emission_factor(sc.tex, lmid, lg, lbary));
              a_tmax = fmaxr(dist * (1.0f - 1.0e-4f), 0.0f);
              a_skip = lg;
              a_have = !is_black(a_le);
            }
          }
        }
      }
    }
    MatSample ms;
    load_mat(*M, sn, ms);
    if (TEX && alive) {
      uint2 run = sc.tex.mat ? sc.tex.mat[mid] : make_uint2(0u, 0u);
      uint32_t attr = sc.tex.attr ? sc.tex.attr[g] : 0u;
      if (run.y || (attr & TA_COLORS) || (M->aniso > 0.0f && (attr & TA_TANGENTS))) {
        bool flip = !front && (M->flags & MAT_DS);
        TexCone cone{0.0f, d};
        if (k == 0) cone.width = fr.projection == 1u ? fr.spread : fr.spread * t;  // src/pathtrace.rs:2083-2092
        tex_apply(sc.tex, run, g, attr, bary, sc.pos + 9ull * g, cone, flip ? neg(sn) : sn, flip, *M, ms);
      }
    }
    const float thickness = M->thickness, transmission = EXT ? M->transmission : 0.0f;
    if (EXT && alive && M->dispersion > 0.0f && thickness > 0.0f && transmission > 0.0f) {
      uint32_t c;
      if (hero != 0xffu) {
        c = hero;
      } else {
        uint32_t path_seed = hash_mix(ps, index);
        float cf = coin(path_seed, (k << 16) | 0x7fffu, 0xffffffffu) * 3.0f;
        c = cf > 0.0f ? (uint32_t)cf : 0u;
        c = c > 2u ? 2u : c;
        beta = v3(c == 0 ? beta.x * 3.0f : 0.0f, c == 1 ? beta.y * 3.0f : 0.0f, c == 2 ? beta.z * 3.0f : 0.0f);
        hero = c;
      }
      ms.ior = dispersion_ior(ms.ior, M->dispersion, (int)c);
    }
    BsdfT<EXT> bsdf;
    bsdf.init(ms, neg(d), ngeo, front);
    if (alive && bsdf.black()) alive = false;
    V3 ng = bsdf.ng;
    // One-sample MIS with the guide at eligible vertices: the first kmax
    // bounces, no transmission, not a near-mirror metal (the BSDF sampler
    // alone serves those).
    float gmix = 0.0f;
    const GuideLobes* GL = nullptr;
    if (GUIDE) {
      GL = gu.lobes + i;
      if (alive && k < gu.kmax && transmission == 0.0f && (!EXT || M->dt == 0.0f) &&
          !(M->metallic >= 0.99f && M->roughness < 0.2f))
        gmix = gu.mix;
    }
    // NEE: punctual lights, one shadow-ray slot each.
    bool any_shadow = false;
    // Uniform trip count (nslots - 1 covers every scene of the batch) so
    // every lane reaches the appends.
    for (uint32_t j = 0; PUNCT && j + 1u < nslots; ++j) {
      bool emit = false;
      V3 o2, l, contrib;
      float tmax = 0.0f;
      if (alive && j < sc.num_lights) {
        const DevLight& L = sc.lights[j];
        V3 lrad;
        bool ok = true;
        if (L.kind == 0u) {
          l = neg(ld3(L.dir));
          tmax = __builtin_huge_valf();
          lrad = ld3(L.base);
        } else {
          V3 to = sub(ld3(L.pos), p);
          float d2 = dot(to, to);
          ok = !(isnan(d2) || d2 <= 1.0e-12f);
          float dist = sqrtf(d2);
          l = v3(to.x / dist, to.y / dist, to.z / dist);
          float att = 1.0f / d2;
          if (L.has_range) {
            float x = dist / L.range;
            att *= clampr(1.0f - x * x * x * x, 0.0f, 1.0f);
          }
          if (L.kind == 2u) {
            float cd = dot(ld3(L.dir), neg(l));
            float a = clampr(cd * L.spot_scale + L.spot_offset, 0.0f, 1.0f);
            att *= a * a;
          }
          ok = ok && att > 0.0f;
          lrad = v3(L.base[0] * att, L.base[1] * att, L.base[2] * att);
          tmax = isfinite(dist) ? dist * (1.0f - 1.0e-4f) : __builtin_huge_valf();
        }
        if (ok) {
          V3 f;
          float pdf_unused;
          bsdf.eval(l, f, pdf_unused);
          if (!is_black(f)) {
            o2 = spawn(p, ng, l);
            contrib = clampc(mul(mul(beta, f), lrad), fr.clamp);
            emit = true;
          }
        }
      }
      const float4 ra = make_float4(o2.x, o2.y, o2.z, __uint_as_float(NO_SKIP));
      const float4 rb = make_float4(l.x, l.y, l.z, tmax);
      const size_t slot = (size_t)pid * nslots + j;
      if (stage) {
        stage_entry(s_sa, s_sb, s_si, wbase, staged, emit, ra, rb, (uint32_t)slot);
      } else {
        uint32_t pos = wave_append(scount, emit);
        if (emit) {
          sray[2 * pos] = ra;
          sray[2 * pos + 1] = rb;
          sidx[pos] = (uint32_t)slot;
        }
      }
      if (emit) {
        nee[slot] = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
        vis[slot] = 1;
        any_shadow = true;
      }
    }
    // NEE: the area set's sample drawn above, now with the BSDF
    // (src/pathtrace.rs:2173-2279).
    {
      bool emit = false;
      V3 o2, contrib;
      const V3 l = a_l;
      const float tmax = a_tmax;
      const uint32_t skip = a_skip;
      if (AREA && alive && a_have) {
        V3 f;
        float pb;
        bsdf.eval(l, f, pb);
        if (!is_black(f)) {
          if (GUIDE && gmix > 0.0f) pb = gmix * guide_pdf(GL, l) + (1.0f - gmix) * pb;
          float w = fr.strategy == 0u ? power_heuristic(a_pl, pb) : 1.0f;
          o2 = spawn(p, ng, l);
          contrib = clampc(scale(mul(mul(beta, f), a_le), w / a_pl), fr.clamp);
          emit = true;
        }
      }
      const float4 ra = make_float4(o2.x, o2.y, o2.z, __uint_as_float(skip));
      const float4 rb = make_float4(l.x, l.y, l.z, tmax);
      const size_t slot = (size_t)pid * nslots + (nslots - 1);
      if (stage) {
        stage_entry(s_sa, s_sb, s_si, wbase, staged, emit, ra, rb, (uint32_t)slot);
      } else {
        uint32_t pos = wave_append(scount, emit);
        if (emit) {
          sray[2 * pos] = ra;
          sray[2 * pos + 1] = rb;
          sidx[pos] = (uint32_t)slot;
        }
      }
      if (emit) {
        nee[slot] = make_float4(contrib.x, contrib.y, contrib.z, 0.0f);
        vis[slot] = 1;
        any_shadow = true;
      }
    }
    if (any_shadow) flags |= PF_PENDING;
    // BSDF sampling, medium transitions, roulette (src/pathtrace.rs:2281-2322).
    bool cont = false;
    V3 no, nl;
    if (alive) {
      F4 pb_u = pattern(ps, index, 1 + 2 * k);
      V3 l, f;
      float pdf;
      bool ok;
      if (GUIDE && gmix > 0.0f) {
        // Direction from the guide with probability gmix, else from the
        // BSDF; weighted by the mixture pdf of both.
        F4 gq = pattern(ps, index, G_PATTERN + k);
        float pbs;
        if (gq.v[0] < gmix) {
          l = guide_sample(GL, gq.v[1], gq.v[2], gq.v[3]);
          bsdf.eval(l, f, pbs);
          ok = !is_black(f);
        } else {
          ok = bsdf.sample(pb_u.v[0], pb_u.v[1], pb_u.v[2], l, f, pbs);
        }
        if (ok) {
          pdf = gmix * guide_pdf(GL, l) + (1.0f - gmix) * pbs;
          ok = pdf > 0.0f && isfinite(pdf);
          // Roulette sees the throughput the BSDF sampler alone would give:
          // a guided direction into a bright region has a small f / p by
          // design, and throughput roulette would end most such paths and
          // scale the survivors up, undoing the guide.
          rr_scale = pbs > 0.0f ? fminf(rr_scale * (pdf / pbs), 1.0e30f) : 1.0e30f;
        }
      } else {
        ok = bsdf.sample(pb_u.v[0], pb_u.v[1], pb_u.v[2], l, f, pdf);
      }
      if (ok) {
        beta = mul(beta, scale(f, 1.0f / pdf));
        if (finite3(beta)) {
          prev_pdf = pdf;
          prev_pos = p;
          if (dot(l, ng) < 0.0f && thickness > 0.0f && transmission > 0.0f) {
            if (front) {
              float dist = M->att_dist;
              if (isfinite(dist)) {
                sg = make_float4(-logf(fmaxr(M->att_color[0], 1.0e-6f)) / dist,
                                 -logf(fmaxr(M->att_color[1], 1.0e-6f)) / dist,
                                 -logf(fmaxr(M->att_color[2], 1.0e-6f)) / dist, 1.0f);
              } else {
                sg = make_float4(0, 0, 0, 0);
              }
            } else {
              sg = make_float4(0, 0, 0, 0);
            }
          }
          cont = true;
          if (k + 1 >= fr.rr_start) {
            float q = clampr(GUIDE ? max3(beta) * rr_scale : max3(beta), 0.05f, 1.0f);
            if (q < 1.0f) {
              if (pb_u.v[3] >= q)
                cont = false;
              else
                beta = scale(beta, 1.0f / q);
            }
          }
          if (cont) {
            no = spawn(p, ng, l);
            nl = l;
          }
        }
      }
    }
    if (GUIDE && gu.rec && cont && gmix > 0.0f && k < (uint32_t)G_KREC && (pid & rec_mask) == 0u) {
      float4* r = gu.rec + ((size_t)(pid >> gu.shift) * G_KREC + k) * 4;
      r[0] = make_float4(p.x, p.y, p.z, prev_pdf);
      r[1] = make_float4(nl.x, nl.y, nl.z, 1.0f);
      r[2] = make_float4(beta.x, beta.y, beta.z, 0.0f);
    }
    {
      const float4 na = make_float4(no.x, no.y, no.z, __uint_as_float(k + 1));
      const float4 nb = make_float4(nl.x, nl.y, nl.z, __builtin_huge_valf());
      if (stage) {
        stage_entry(s_na, s_nb, s_np, nbase, nstaged, cont, na, nb, pid);
      } else {
        uint32_t pos = wave_append(ncount, cont);
        if (cont) {
          nray[2 * pos] = na;
          nray[2 * pos + 1] = nb;
          npid[pos] = pid;
        }
      }
    }
    if (valid) {
      if (!cont) flags |= PF_DONE;
      beta_s[pid] = make_float4(beta.x, beta.y, beta.z, prev_pdf);
      rad_s[pid] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
      prev_s[pid] = make_float4(prev_pos.x, prev_pos.y, prev_pos.z, GUIDE ? rr_scale : 0.0f);
      if (EXT) sigma_s[pid] = sg;
      meta_s[pid] = make_uint4(meta.x, meta.y, ((k + (cont ? 1u : 0u)) & 0xffffu) | (hero << 16) | (flags << 24), ps);
    }
    if (stage)
      block_flush(s_cnt, s_base, s_sa, s_sb, s_si, wbase, staged, scount, sray, sidx, s_na, s_nb, s_np, nbase, nstaged,
                  ncount, nray, npid);
  }
}

// Guide training samples of the recorded paths, once their radiance is
// final: incident radiance estimate L_i = (L_final - L_mark) / beta per
// channel, target = luminance(max(L_i, 0)), weight = target / pdf.
// out: each path's final radiance, by path (k_finalize's copy in the path
// radiance state).
__global__ void k_guide_collect(uint32_t P, uint32_t nsel, uint32_t shift, const float4* __restrict__ rec,
                                const float4* __restrict__ out, float4* __restrict__ samp, uint32_t* count,
                                uint32_t cap) {
  uint32_t sp = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t pid = sp << shift;
  bool live = sp < nsel && pid < P;
  float4 fin = live ? out[pid] : make_float4(0, 0, 0, 0);
  for (int k = 0; k < G_KREC; ++k) {
    bool emit = false;
    float4 r0, r1;
    float w = 0.0f;
    if (live) {
      const float4* r = rec + ((size_t)sp * G_KREC + k) * 4;
      r1 = r[1];
      if (r1.w == 1.0f) {
        r0 = r[0];
        float4 b = r[2], lm = r[3];
        float li[3] = {b.x > 0.0f ? (fin.x - lm.x) / b.x : 0.0f, b.y > 0.0f ? (fin.y - lm.y) / b.y : 0.0f,
                       b.z > 0.0f ? (fin.z - lm.z) / b.z : 0.0f};
        float t = luminance(v3(fmaxr(li[0], 0.0f), fmaxr(li[1], 0.0f), fmaxr(li[2], 0.0f)));
        w = t / r0.w;
        emit = isfinite(w) && w > 0.0f;
      }
    }
    uint32_t pos = wave_append(count, emit);
    if (emit && pos < cap) {
      samp[2ull * pos] = make_float4(r0.x, r0.y, r0.z, w);
      samp[2ull * pos + 1] = make_float4(r1.x, r1.y, r1.z, r0.w);
    }
  }
}

// ---------------------------------------------------------------------
// Parity entry points of the texture layer: PreparedTexture::sample_lod
// on host (u, v, lod) triples, and material_oriented + emission + the
// any-hit alpha at host (triangle, barycentrics) hits, 43 floats each in
// the layout of the oracle's material evaluation.
// ---------------------------------------------------------------------
__global__ void k_tex_sample(DevTextures T, uint32_t chain, uint32_t n, const float* __restrict__ uv,
                             const float* __restrict__ lod, float4* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  F4 r = tex_sample_lod(T, chain, uv[2 * i], uv[2 * i + 1], lod[i]);
  out[i] = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
}

__global__ void k_material_eval(DevScene sc, uint32_t n, const uint32_t* __restrict__ gids,
                                const float* __restrict__ bary3, const float* __restrict__ width,
                                const float* __restrict__ dir3, const uint8_t* __restrict__ flips,
                                float* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t g = gids[i];
  V3 bary = v3(bary3[3 * i], bary3[3 * i + 1], bary3[3 * i + 2]);
  const float* tv = sc.pos + 9ull * g;
  V3 p0 = ld3(tv), p1 = ld3(tv + 3), p2 = ld3(tv + 6);
  V3 cr = cross(sub(p1, p0), sub(p2, p0));
  float len = sqrtf(dot(cr, cr));
  V3 sn = len > 0.0f ? v3(cr.x / len, cr.y / len, cr.z / len) : v3(0.0f, 0.0f, 1.0f);
  uint32_t info = sc.tri_info[g];
  uint32_t mid = info & TI_MAT_MASK;
  if (info & TI_HAS_NORMALS) {
    const float* nv = sc.nrm + 9ull * g;
    V3 n0 = ld3(nv), n1 = ld3(nv + 3), n2 = ld3(nv + 6);
    V3 nn = normalise(v3(n0.x * bary.x + n1.x * bary.y + n2.x * bary.z, n0.y * bary.x + n1.y * bary.y + n2.y * bary.z,
                         n0.z * bary.x + n1.z * bary.y + n2.z * bary.z));
    sn = dot(nn, nn) > 0.5f ? nn : sn;
  }
  const DevMaterial& M = sc.mats[mid];
  bool flip = flips[i] != 0;
  MatSample ms;
  load_mat(M, sn, ms);
  TexCone cone{width[i], v3(dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2])};
  uint2 run = sc.tex.mat ? sc.tex.mat[mid] : make_uint2(0u, 0u);
  uint32_t attr = sc.tex.attr ? sc.tex.attr[g] : 0u;
  float base_a = M.base[3];
  V3 emissive = ld3(M.emissive);
  if (run.y || (attr & TA_COLORS) || (M.aniso > 0.0f && (attr & TA_TANGENTS))) {
    tex_apply(sc.tex, run, g, attr, bary, tv, cone, flip ? neg(sn) : sn, flip, M, ms);
    F4 t;
    const DevTexBind* b = tex_find(sc.tex, run, TS_BASE);
    if (b && tex_tap(sc.tex, *b, g, attr, bary, tv, cone, t)) base_a *= t.v[3];
    if (attr & TA_COLORS) base_a *= tex_interp4(sc.tex.col + 12ull * g, bary).v[3];
    b = tex_find(sc.tex, run, TS_EMISSIVE);
    if (b && tex_tap(sc.tex, *b, g, attr, bary, tv, cone, t)) emissive = mul(emissive, v3(t.v[0], t.v[1], t.v[2]));
  } else if (flip) {
    ms.normal = sn;  // flipped and flipped back
    ms.clearcoat_normal = sn;
  }
  V3 le = ((M.flags & MAT_UNLIT) || is_black(ld3(M.emissive))) ? v3(0, 0, 0)
                                                               : mul(ld3(M.emissive), tex_emission_factor(sc.tex, mid, g, bary));
  float alpha = tex_alpha(sc.tex, M.base[3], mid, g, bary);
  float v[43] = {ms.base.x, ms.base.y, ms.base.z, ms.metallic, ms.roughness, ms.normal.x, ms.normal.y, ms.normal.z,
                 ms.ior, ms.specular, ms.specular_color.x, ms.specular_color.y, ms.specular_color.z, ms.transmission,
                 ms.thickness, ms.clearcoat, ms.clearcoat_roughness, ms.clearcoat_normal.x, ms.clearcoat_normal.y,
                 ms.clearcoat_normal.z, ms.sheen_color.x, ms.sheen_color.y, ms.sheen_color.z, ms.sheen_roughness,
                 ms.anisotropy, ms.anisotropy_direction.x, ms.anisotropy_direction.y, ms.anisotropy_direction.z,
                 ms.iridescence, ms.iridescence_ior, ms.iridescence_thickness, ms.diffuse_transmission,
                 ms.diffuse_transmission_color.x, ms.diffuse_transmission_color.y, ms.diffuse_transmission_color.z,
                 base_a, emissive.x, emissive.y, emissive.z, le.x, le.y, le.z, alpha};
  for (int k = 0; k < 43; ++k) out[43ull * i + k] = v[k];
}

// Per-sample radiance: the pending NEE of the last vertex, then the
// finite check of Integrator::sample (src/pathtrace.rs:2324-2328).
// The path's final radiance: in rad_s (by path, for the guide's training
// samples) and in out at its segment's s * npx + pixel offset (k_reduce).
__global__ void k_finalize(const DevJob* __restrict__ jobs, SegTable segs, uint32_t P, uint32_t nslots,
                           const float4* __restrict__ nee, uint8_t* __restrict__ vis, float4* __restrict__ rad_s,
                           const uint4* __restrict__ meta_s, float4* __restrict__ out) {
  uint32_t pid = blockIdx.x * blockDim.x + threadIdx.x;
  if (pid >= P) return;
  uint4 meta = meta_s[pid];
  uint32_t flags = meta.z >> 24;
  V3 radiance = f4xyz(rad_s[pid]);
  if (flags & PF_PENDING) add_pending(pid, nslots, nee, vis, radiance);
  if (!finite3(radiance)) radiance = v3(0.0f, 0.0f, 0.0f);
  const Seg sg = segs.s[seg_of(segs, pid)];
  uint32_t s, off;
  path_sample(sg, jobs[sg.job].fr.win_w, pid - sg.path0, s, off);
  rad_s[pid] = make_float4(radiance.x, radiance.y, radiance.z, 0.0f);
  out[sg.path0 + (size_t)s * sg.npx + off] =
      make_float4(radiance.x, radiance.y, radiance.z, (flags & PF_UNCOVERED) ? 0.0f : 1.0f);
}

// Accumulate S samples per pixel in sample order into the f64
// accumulator (src/pathtrace.rs:2569-2578) and the squared radiance.
__global__ void k_reduce(uint32_t npx, uint32_t S, const float4* __restrict__ out, double* __restrict__ accum,
                         double* __restrict__ sumsq) {
  uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  if (px >= npx) return;
  double a0 = accum[4ull * px], a1 = accum[4ull * px + 1], a2 = accum[4ull * px + 2], a3 = accum[4ull * px + 3];
  double q0 = sumsq[3ull * px], q1 = sumsq[3ull * px + 1], q2 = sumsq[3ull * px + 2];
  for (uint32_t s = 0; s < S; ++s) {
    float4 c = out[(size_t)s * npx + px];
    if (c.w > 0.0f) {
      double r = (double)c.x, g = (double)c.y, b = (double)c.z;
      a0 += r;
      a1 += g;
      a2 += b;
      a3 += 1.0;
      q0 += r * r;
      q1 += g * g;
      q2 += b * b;
    }
  }
  accum[4ull * px] = a0;
  accum[4ull * px + 1] = a1;
  accum[4ull * px + 2] = a2;
  accum[4ull * px + 3] = a3;
  sumsq[3ull * px] = q0;
  sumsq[3ull * px + 1] = q1;
  sumsq[3ull * px + 2] = q2;
}

__global__ void k_zero_ctrl(Ctrl* c) {
  if (threadIdx.x == 0) {
    c->qcount[0].v = c->qcount[1].v = 0;
    c->scount.v = 0;
    c->error = 0;
    c->rays_closest = 0;
    c->rays_shadow = 0;
    c->trav[0].v = c->trav[1].v = c->trav[2].v = c->trav[3].v = 0;
  }
  if (threadIdx.x < 16) c->work[threadIdx.x].v = 0;
}

__global__ void k_fill_f64(double* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------
// Resolve (PathTracer::hdr / image, src/pathtrace.rs:2593-2734;
// ToneMap::apply, linear_to_srgb_byte, src/hdr.rs:44-129).
// ---------------------------------------------------------------------
struct DevDisplay {
  float bg[4];
  uint8_t bg_bytes[4];
  uint32_t tone_map;
  float exposure;
};

OXR_D float sanitize(float v) { return isnan(v) ? 0.0f : v; }
OXR_D float aces(float x) {
  x = x * 0.6f;
  float v = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
  return clampr(v, 0.0f, 1.0f);
}
OXR_D V3 tone_map(uint32_t tm, V3 rgb) {
  rgb = v3(fmaxr(sanitize(rgb.x), 0.0f), fmaxr(sanitize(rgb.y), 0.0f), fmaxr(sanitize(rgb.z), 0.0f));
  if (tm == 1u) {
    float l = 0.2126f * rgb.x + 0.7152f * rgb.y + 0.0722f * rgb.z;
    if (l <= 0.0f) return v3(0, 0, 0);
    float s = 1.0f / (1.0f + l);
    return v3(fminr(rgb.x * s, 1.0f), fminr(rgb.y * s, 1.0f), fminr(rgb.z * s, 1.0f));
  }
  if (tm == 2u) return v3(aces(rgb.x), aces(rgb.y), aces(rgb.z));
  return v3(fminr(rgb.x, 1.0f), fminr(rgb.y, 1.0f), fminr(rgb.z, 1.0f));
}
OXR_D uint8_t srgb_byte(float c) {
  if (isnan(c)) return 0;
  int lo = 0, hi = 255;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (c_srgb_th[mid] <= c)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (uint8_t)lo;
}

OXR_D bool resolve_px(const double* a, uint32_t samples, const DevDisplay& dd, bool map, float o[4]) {
  if (a[3] <= 0.0 || samples == 0) return false;
  V3 m = v3((float)(a[0] / a[3]), (float)(a[1] / a[3]), (float)(a[2] / a[3]));
  float f = (float)clampd(a[3] / (double)samples, 0.0, 1.0);
  V3 c = map ? tone_map(dd.tone_map, scale(m, dd.exposure)) : m;
  if (f >= 1.0f) {
    o[0] = c.x;
    o[1] = c.y;
    o[2] = c.z;
    o[3] = 1.0f;
    return true;
  }
  float a_sum = f + (1.0f - f) * dd.bg[3];
  float cc[3] = {c.x, c.y, c.z};
  if (a_sum > 0.0f) {
    for (int k = 0; k < 3; ++k) o[k] = (f * cc[k] + (1.0f - f) * dd.bg[k] * dd.bg[3]) / a_sum;
    o[3] = a_sum;
  } else {
    for (int k = 0; k < 3; ++k) o[k] = f * cc[k] + (1.0f - f) * dd.bg[k];
    o[3] = 0.0f;
  }
  return true;
}

__global__ void k_resolve_hdr(uint32_t npx, uint32_t samples, const double* __restrict__ accum, DevDisplay dd,
                              float4* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npx) return;
  float o[4];
  if (!resolve_px(accum + 4ull * i, samples, dd, false, o)) {
    out[i] = make_float4(dd.bg[0], dd.bg[1], dd.bg[2], dd.bg[3]);
    return;
  }
  out[i] = make_float4(o[0], o[1], o[2], o[3]);
}

__global__ void k_resolve_rgba8(uint32_t npx, uint32_t samples, const double* __restrict__ accum, DevDisplay dd,
                                uchar4* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npx) return;
  float o[4];
  if (!resolve_px(accum + 4ull * i, samples, dd, true, o)) {
    out[i] = make_uchar4(dd.bg_bytes[0], dd.bg_bytes[1], dd.bg_bytes[2], dd.bg_bytes[3]);
    return;
  }
  float a = clampr(o[3], 0.0f, 1.0f) * 255.0f;
  out[i] = make_uchar4(srgb_byte(o[0]), srgb_byte(o[1]), srgb_byte(o[2]), (uint8_t)roundf(a));
}

// ---------------------------------------------------------------------
// Parity kernels: Sobol' points and the BSDF.
// ---------------------------------------------------------------------
__global__ void k_sobol(uint32_t n, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ seed,
                        float4* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  F4 p = sobol_owen_4d(idx[i], seed[i]);
  out[i] = make_float4(p.v[0], p.v[1], p.v[2], p.v[3]);
}

OXR_D void ms_from_abi(const float* m, MatSample& ms) {
  // oxr_material_sample field order (include/oxr/pathtrace.h).
  ms.base = ld3(m + 0);
  ms.metallic = m[3];
  ms.roughness = m[4];
  ms.normal = ld3(m + 5);
  ms.ior = m[8];
  ms.specular = m[9];
  ms.specular_color = ld3(m + 10);
  ms.transmission = m[13];
  ms.thickness = m[14];
  ms.clearcoat = m[15];
  ms.clearcoat_roughness = m[16];
  ms.clearcoat_normal = ld3(m + 17);
  ms.sheen_color = ld3(m + 20);
  ms.sheen_roughness = m[23];
  ms.anisotropy = m[24];
  ms.anisotropy_direction = ld3(m + 25);
  ms.iridescence = m[28];
  ms.iridescence_ior = m[29];
  ms.iridescence_thickness = m[30];
  ms.diffuse_transmission = m[31];
  ms.diffuse_transmission_color = ld3(m + 32);
}

// Environment map evaluation / sampling on host inputs (parity entry
// points): in = directions (eval) or (u0, u1) pairs (sample).
__global__ void k_env(DevEnv e, uint32_t n, int do_sample, const float* __restrict__ in, float* __restrict__ dirs,
                      float* __restrict__ pdf, float* __restrict__ le, uint8_t* __restrict__ ok) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3 d, r = v3(0.0f, 0.0f, 0.0f);
  float p = 0.0f;
  if (do_sample) {
    bool got = env_sample(e, in[2 * i], in[2 * i + 1], d, r, p);
    ok[i] = got ? 1 : 0;
    dirs[3 * i] = d.x;
    dirs[3 * i + 1] = d.y;
    dirs[3 * i + 2] = d.z;
  } else {
    d = ld3(in + 3 * i);
    p = env_pdf(e, d);
    r = env_radiance(e, d);
  }
  pdf[i] = p;
  le[3 * i] = r.x;
  le[3 * i + 1] = r.y;
  le[3 * i + 2] = r.z;
}

__global__ void k_bsdf(uint32_t n, int do_sample, const float* __restrict__ ms_abi, const float* __restrict__ v,
                       const float* __restrict__ ng, const uint8_t* __restrict__ front, const float* __restrict__ in,
                       float* __restrict__ l_out, float* __restrict__ f_out, float* __restrict__ pdf_out,
                       uint8_t* __restrict__ ok_out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  MatSample ms;
  ms_from_abi(ms_abi + 35ull * i, ms);
  Bsdf b;
  b.init(ms, ld3(v + 3ull * i), ld3(ng + 3ull * i), front[i] != 0);
  V3 f = v3(0, 0, 0), l = v3(0, 0, 0);
  float pdf = 0.0f;
  if (do_sample) {
    bool ok = b.sample(in[3 * i], in[3 * i + 1], in[3 * i + 2], l, f, pdf);
    if (!ok) {
      l = v3(0, 0, 0);
      f = v3(0, 0, 0);
      pdf = 0.0f;
    }
    ok_out[i] = ok ? 1 : 0;
    l_out[3 * i] = l.x;
    l_out[3 * i + 1] = l.y;
    l_out[3 * i + 2] = l.z;
  } else {
    b.eval(ld3(in + 3ull * i), f, pdf);
  }
  f_out[3 * i] = f.x;
  f_out[3 * i + 1] = f.y;
  f_out[3 * i + 2] = f.z;
  pdf_out[i] = pdf;
}

// ---------------------------------------------------------------------
// Host-side launchers (declared in runtime.h).
// ---------------------------------------------------------------------
namespace launch {

hipError_t upload_constants(const uint32_t* sobol128, const float* sheen512, const float* srgb255) {
  static uint32_t tab[3][4][256];
  for (int d = 1; d < 4; ++d)
    for (int by = 0; by < 4; ++by)
      for (int v = 0; v < 256; ++v) {
        uint32_t x = 0;
        for (int j = 0; j < 8; ++j)
          if ((v >> j) & 1) x ^= sobol128[d * 32 + by * 8 + j];
        tab[d - 1][by][v] = x;
      }
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_sobol_tab), tab, sizeof(tab));
  if (e != hipSuccess) return e;
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_sheen), sheen512, sizeof(float) * 512);
  if (e != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_srgb_th), srgb255, sizeof(float) * 255);
}

int trace_blocks_per_cu(bool any, bool alpha) {
  // The persistent grid must fit on the device for the single- and the
  // mixed-scene variant of the kernel (the textured-alpha variants hold
  // more registers and get their own grid).
  int nb[2] = {0, 0};
  hipError_t e[2];
  if (any && alpha) {
    e[0] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[0], k_anyhit<true, false, true>, TRACE_TB, 0);
    e[1] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[1], k_anyhit<true, true, true>, TRACE_TB, 0);
  } else if (any) {
    e[0] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[0], k_anyhit<true, false, false>, TRACE_TB, 0);
    e[1] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[1], k_anyhit<true, true, false>, TRACE_TB, 0);
  } else if (alpha) {
    e[0] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[0], k_closest<true, false, true>, TRACE_TB, 0);
    e[1] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[1], k_closest<true, true, true>, TRACE_TB, 0);
  } else {
    e[0] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[0], k_closest<true, false, false>, TRACE_TB, 0);
    e[1] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb[1], k_closest<true, true, false>, TRACE_TB, 0);
  }
  if (e[0] != hipSuccess || e[1] != hipSuccess || nb[0] <= 0 || nb[1] <= 0) return 4;
  return std::min(nb[0], nb[1]);
}

int trace_block_threads() { return TRACE_TB; }
int stack_lds_entries() { return STACK; }

void closest(hipStream_t s, int grid, const DevScene& sc, const DevJob* jobs, const uint32_t* pjob, bool path,
             bool multi, const float4* rays, const uint32_t* qpid, const uint4* meta, uint32_t fixed_count,
             const uint32_t* count_ptr, uint32_t* work, uint32_t cull_all, uint4* hits, uint32_t* spill,
             uint32_t spill_cap, uint32_t* err, uint32_t nseg, const DevScene* asc, unsigned long long* counts) {
  {
    uint32_t ns = nseg & 0xffu, rf = (nseg >> 8) & 0xffu;
    if (ns < 1 || ns > 8 || grid % (int)ns != 0) ns = 1;
    if (rf == 0 || rf > 64) rf = 64;
    nseg = ns | (rf << 8);
  }
  // asc (the scene in device memory) selects the textured-alpha variant;
  // a mixed batch finds each path's scene through its job instead.
  const bool alpha = asc != nullptr;
#define OXR_CLOSEST(P, M, A)                                                                                     \
  hipLaunchKernelGGL((k_closest<P, M, A>), dim3(grid), dim3(TRACE_TB), 0, s, sc, jobs, pjob, rays, qpid, meta,      \
                     fixed_count, count_ptr, work, cull_all, hits, spill, spill_cap, err, nseg, asc, nullptr)
  if (counts && path && !multi && !alpha)
    hipLaunchKernelGGL((k_closest<true, false, false, true>), dim3(grid), dim3(TRACE_TB), 0, s, sc, jobs, pjob, rays,
                       qpid, meta, fixed_count, count_ptr, work, cull_all, hits, spill, spill_cap, err, nseg, asc,
                       counts);
  else if (!path && alpha)
    OXR_CLOSEST(false, false, true);
  else if (!path)
    OXR_CLOSEST(false, false, false);
  else if (multi && alpha)
    OXR_CLOSEST(true, true, true);
  else if (multi)
    OXR_CLOSEST(true, true, false);
  else if (alpha)
    OXR_CLOSEST(true, false, true);
  else
    OXR_CLOSEST(true, false, false);
#undef OXR_CLOSEST
}

void anyhit(hipStream_t s, int grid, const DevScene& sc, const DevJob* jobs, const uint32_t* pjob, bool path,
            bool multi, const float4* rays, const uint32_t* sidx, const uint4* meta, uint32_t nslots,
            uint32_t fixed_count, const uint32_t* count_ptr, uint32_t* work, uint8_t* out, uint32_t* spill,
            uint32_t spill_cap, uint32_t* err, uint32_t nseg, const DevScene* asc, unsigned long long* counts) {
  {
    uint32_t ns = nseg & 0xffu, rf = (nseg >> 8) & 0xffu;
    if (ns < 1 || ns > 8 || grid % (int)ns != 0) ns = 1;
    if (rf == 0 || rf > 64) rf = 64;
    nseg = ns | (rf << 8);
  }
  const bool alpha = asc != nullptr;
#define OXR_ANYHIT(P, M, A)                                                                                      \
  hipLaunchKernelGGL((k_anyhit<P, M, A>), dim3(grid), dim3(TRACE_TB), 0, s, sc, jobs, pjob, rays, sidx, meta,      \
                     nslots, fixed_count, count_ptr, work, out, spill, spill_cap, err, nseg, asc, nullptr)
  if (counts && path && !multi && !alpha)
    hipLaunchKernelGGL((k_anyhit<true, false, false, true>), dim3(grid), dim3(TRACE_TB), 0, s, sc, jobs, pjob, rays,
                       sidx, meta, nslots, fixed_count, count_ptr, work, out, spill, spill_cap, err, nseg, asc, counts);
  else if (!path && alpha)
    OXR_ANYHIT(false, false, true);
  else if (!path)
    OXR_ANYHIT(false, false, false);
  else if (multi && alpha)
    OXR_ANYHIT(true, true, true);
  else if (multi)
    OXR_ANYHIT(true, true, false);
  else if (alpha)
    OXR_ANYHIT(true, false, true);
  else
    OXR_ANYHIT(true, false, false);
#undef OXR_ANYHIT
}

void camera(hipStream_t s, const DevJob* jobs, const SegTable& segs, uint32_t P, const PathBuffers& pb, Ctrl* ctrl) {
  hipLaunchKernelGGL(k_camera, dim3((P + 255) / 256), dim3(256), 0, s, jobs, segs, P, pb.qray[0], pb.qpid[0],
                     pb.meta, pb.pjob, ctrl);
}

void prep(hipStream_t s, Ctrl* ctrl, uint32_t cur, uint32_t count_shadow_prev) {
  hipLaunchKernelGGL(k_prep, dim3(1), dim3(64), 0, s, ctrl, cur, count_shadow_prev);
}

void shade(hipStream_t s, int grid, int waves, const DevJob* jobs, uint32_t nslots, const PathBuffers& pb, Ctrl* ctrl,
           uint32_t cur, const DevGuideUse* guide, bool tex, bool multi, int ext, bool area, uint32_t kb) {
  DevGuideUse gu = guide ? *guide : DevGuideUse{nullptr, nullptr, 0.0f, 0u, 0u};
#define OXR_SHADE_P(W, G, T, M, X, A, P)                                                                         \
  hipLaunchKernelGGL((k_shade<W, G, T, M, X, A, P>), dim3(grid), dim3(256), 0, s, jobs, pb.pjob, nslots, pb.qray[cur], \
                     pb.qpid[cur], pb.hits, pb.qray[cur ^ 1u], pb.qpid[cur ^ 1u], pb.sray, pb.sidx, pb.beta, pb.rad, \
                     pb.prev, pb.sigma, pb.meta, pb.nee, pb.vis, ctrl, cur, gu, kb)
#define OXR_SHADE(W, G, T, M, X, A) OXR_SHADE_P(W, G, T, M, X, A, true)
  // Mixed batches (the server), textured scenes and guided scenes with
  // punctual lights run at one wave per SIMD; the occupancy variants are
  // for single-scene shading (guided: lit by the area set only). Plain and
  // mixed batches have variants without the extension lobes (or without
  // iridescence) and without area lights.
#define OXR_SHADE_XA(W, M)                                 \
  do {                                                     \
    if (ext >= 2 && area)                                  \
      OXR_SHADE(W, false, false, M, 2, true);              \
    else if (ext >= 2)                                     \
      OXR_SHADE(W, false, false, M, 2, false);             \
    else if (ext == 1 && area)                             \
      OXR_SHADE(W, false, false, M, 1, true);              \
    else if (ext == 1)                                     \
      OXR_SHADE(W, false, false, M, 1, false);             \
    else if (area)                                         \
      OXR_SHADE(W, false, false, M, 0, true);              \
    else                                                   \
      OXR_SHADE(W, false, false, M, 0, false);             \
  } while (0)
  // Guided single-scene batches lit only by the area set, without
  // iridescence (two waves per SIMD spill 28 bytes a lane at EXT = 1).
#define OXR_SHADE_GUIDE_AREA_ONLY(W)                       \
  do {                                                     \
    if (ext == 1)                                          \
      OXR_SHADE_P(W, true, false, false, 1, true, false);  \
    else                                                   \
      OXR_SHADE_P(W, true, false, false, 0, true, false);  \
  } while (0)
  // Single-scene batches lit only by the area set.
#define OXR_SHADE_AREA_ONLY(W)                             \
  do {                                                     \
    if (ext >= 2)                                          \
      OXR_SHADE_P(W, false, false, false, 2, true, false); \
    else if (ext == 1)                                     \
      OXR_SHADE_P(W, false, false, false, 1, true, false); \
    else                                                   \
      OXR_SHADE_P(W, false, false, false, 0, true, false); \
  } while (0)
  if (multi && tex)
    OXR_SHADE(1, false, true, true, 2, true);
  else if (multi)
    OXR_SHADE_XA(1, true);
  else if (guide && tex)
    OXR_SHADE(1, true, true, false, 2, true);
  else if (guide && area && nslots == 1u && ext >= 2)
    OXR_SHADE_P(1, true, false, false, 2, true, false);
  else if (guide && area && nslots == 1u && waves >= 2)
    OXR_SHADE_GUIDE_AREA_ONLY(2);
  else if (guide && area && nslots == 1u)
    OXR_SHADE_GUIDE_AREA_ONLY(1);
  else if (guide && ext)
    OXR_SHADE(1, true, false, false, 2, true);
  else if (guide)
    OXR_SHADE(1, true, false, false, 0, true);
  else if (tex && ext)
    OXR_SHADE(1, false, true, false, 2, true);
  else if (tex)
    OXR_SHADE(1, false, true, false, 0, true);
  else if (area && nslots == 1u && waves >= 3)
    OXR_SHADE_AREA_ONLY(3);
  else if (area && nslots == 1u && waves == 2)
    OXR_SHADE_AREA_ONLY(2);
  else if (area && nslots == 1u)
    OXR_SHADE_AREA_ONLY(1);
  else if (waves >= 3)
    OXR_SHADE_XA(3, false);
  else if (waves == 2)
    OXR_SHADE_XA(2, false);
  else
    OXR_SHADE_XA(1, false);
#undef OXR_SHADE_AREA_ONLY
#undef OXR_SHADE_GUIDE_AREA_ONLY
#undef OXR_SHADE_XA
#undef OXR_SHADE
#undef OXR_SHADE_P
}

void guide_collect(hipStream_t s, uint32_t P, const DevGuideUse& gu, const PathBuffers& pb, float4* samp,
                   uint32_t* count, uint32_t cap) {
  uint32_t nsel = (uint32_t)(((uint64_t)P + (1ull << gu.shift) - 1ull) >> gu.shift);
  hipLaunchKernelGGL(k_guide_collect, dim3((nsel + 255) / 256), dim3(256), 0, s, P, nsel, gu.shift, gu.rec, pb.rad,
                     samp, count, cap);
}

void tex_sample(hipStream_t s, const DevTextures& T, uint32_t chain, uint32_t n, const float* uv, const float* lod,
                float4* out) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_tex_sample, dim3((n + 255) / 256), dim3(256), 0, s, T, chain, n, uv, lod, out);
}

void material_eval(hipStream_t s, const DevScene& sc, uint32_t n, const uint32_t* g, const float* bary,
                   const float* width, const float* dir, const uint8_t* flip, float* out) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_material_eval, dim3((n + 127) / 128), dim3(128), 0, s, sc, n, g, bary, width, dir, flip, out);
}

void finalize(hipStream_t s, const DevJob* jobs, const SegTable& segs, uint32_t P, uint32_t nslots,
              const PathBuffers& pb) {
  hipLaunchKernelGGL(k_finalize, dim3((P + 255) / 256), dim3(256), 0, s, jobs, segs, P, nslots, pb.nee, pb.vis, pb.rad,
                     pb.meta, pb.out);
}

void reduce(hipStream_t s, uint32_t npx, uint32_t S, const float4* out, double* accum, double* sumsq) {
  hipLaunchKernelGGL(k_reduce, dim3((npx + 255) / 256), dim3(256), 0, s, npx, S, out, accum, sumsq);
}

void zero_ctrl(hipStream_t s, Ctrl* c) { hipLaunchKernelGGL(k_zero_ctrl, dim3(1), dim3(64), 0, s, c); }

void fill_f64(hipStream_t s, double* p, size_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
}

void resolve(hipStream_t s, bool rgba8, uint32_t npx, uint32_t samples, const double* accum, const float bg[4],
             const uint8_t bgb[4], uint32_t tone_map, float exposure, void* out) {
  DevDisplay dd;
  for (int k = 0; k < 4; ++k) {
    dd.bg[k] = bg[k];
    dd.bg_bytes[k] = bgb[k];
  }
  dd.tone_map = tone_map;
  dd.exposure = exposure;
  if (rgba8)
    hipLaunchKernelGGL(k_resolve_rgba8, dim3((npx + 255) / 256), dim3(256), 0, s, npx, samples, accum, dd,
                       (uchar4*)out);
  else
    hipLaunchKernelGGL(k_resolve_hdr, dim3((npx + 255) / 256), dim3(256), 0, s, npx, samples, accum, dd,
                       (float4*)out);
}

void sobol(hipStream_t s, uint32_t n, const uint32_t* idx, const uint32_t* seed, float4* out) {
  hipLaunchKernelGGL(k_sobol, dim3((n + 255) / 256), dim3(256), 0, s, n, idx, seed, out);
}

void env(hipStream_t s, const DevEnv& e, uint32_t n, int do_sample, const float* in, float* dirs, float* pdf,
         float* le, uint8_t* ok) {
  hipLaunchKernelGGL(k_env, dim3((n + 127) / 128), dim3(128), 0, s, e, n, do_sample, in, dirs, pdf, le, ok);
}

void bsdf(hipStream_t s, uint32_t n, int do_sample, const float* ms, const float* v, const float* ng,
          const uint8_t* front, const float* in, float* l, float* f, float* pdf, uint8_t* ok) {
  hipLaunchKernelGGL(k_bsdf, dim3((n + 127) / 128), dim3(128), 0, s, n, do_sample, ms, v, ng, front, in, l, f, pdf, ok);
}

}  // namespace launch
}  // namespace oxr
