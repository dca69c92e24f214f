"""Host-side BVH builder (oxideav-render_amd/csrc/bvh.cpp), compiled with a
small harness: every triangle lands in exactly one leaf, every leaf box
holds its triangles, the quantised device nodes hold their boxes, and the
tree is the same with 1 and 8 threads —
including sizes where large subtrees are handed to the work queue
(more than 65536 triangles per side) below the multi-thread cut-off, and
nodes big enough (2^20 items) to split their passes over threads."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "oxideav-render_amd", "csrc")

HARNESS = r'''
#include "bvh.h"
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <cmath>
int main(int argc, char** argv) {
  FILE* f = fopen(argv[1], "rb");
  fseek(f, 0, SEEK_END); long sz = ftell(f); fseek(f, 0, SEEK_SET);
  std::vector<float> pos(sz / 4);
  if (fread(pos.data(), 1, sz, f) != (size_t)sz) return 2;
  fclose(f);
  uint32_t n = pos.size() / 9;
  float bmin[3] = {1e30f, 1e30f, 1e30f}, bmax[3] = {-1e30f, -1e30f, -1e30f};
  for (uint32_t i = 0; i < 3 * n; ++i)
    for (int k = 0; k < 3; ++k) {
      bmin[k] = std::fmin(bmin[k], pos[3 * i + k]);
      bmax[k] = std::fmax(bmax[k], pos[3 * i + k]);
    }
  std::vector<uint32_t> flags(n, 0); std::vector<float> alpha(n, 1.0f);
  oxr::Bvh4 b = oxr::build_bvh4(pos.data(), n, flags.data(), alpha.data(), 4, atoi(argv[2]), argc > 4);
  // Leaf containment: every triangle of a leaf lies inside the leaf's box.
  long bad = 0, inleaf = 0;
  for (const auto& nd : b.nodes)
    for (int c = 0; c < 4; ++c) {
      if (nd.child[c] == oxr::NODE_EMPTY || nd.count[c] == 0) continue;
      for (uint32_t k = 0; k < nd.count[c]; ++k) {
        const oxr::TriRec& t = b.tris[nd.child[c] + k];
        const float* v[3] = {t.v0, t.v1, t.v2};
        for (int j = 0; j < 3; ++j)
          if (v[j][0] < nd.lo_x[c] || v[j][0] > nd.hi_x[c] || v[j][1] < nd.lo_y[c] || v[j][1] > nd.hi_y[c] ||
              v[j][2] < nd.lo_z[c] || v[j][2] > nd.hi_z[c]) ++bad;
        ++inleaf;
      }
    }
  // Quantised nodes: same children and counts; every decoded box holds the
  // exact box with a margin of one step (exact arithmetic); and the
  // device's float slab test on them never culls a child that the exact
  // box's test keeps, over random rays.
  std::vector<oxr::Node4Q> q = oxr::quantize_nodes(b.nodes);
  long qbad = 0, culled = 0;
  for (size_t i = 0; i < b.nodes.size(); ++i) {
    const oxr::Node4& nd = b.nodes[i];
    const oxr::Node4Q& nq = q[i];
    const float* lo[3] = {nd.lo_x, nd.lo_y, nd.lo_z};
    const float* hi[3] = {nd.hi_x, nd.hi_y, nd.hi_z};
    const uint32_t wl[3] = {nq.qlo_x, nq.qlo_y, nq.qlo_z}, wh[3] = {nq.qhi_x, nq.qhi_y, nq.qhi_z};
    for (int c = 0; c < 4; ++c) {
      if (nq.child[c] != nd.child[c] || ((nq.counts >> (8 * c)) & 0xffu) != nd.count[c]) ++qbad;
      if (nd.child[c] == oxr::NODE_EMPTY) continue;
      for (int k = 0; k < 3; ++k) {
        const double step = std::ldexp(1.0, (int)(int8_t)((nq.exps >> (8 * k)) & 0xffu));
        const double dl = (double)nq.org[k] + ((wl[k] >> (8 * c)) & 0xffu) * step;
        const double dh = (double)nq.org[k] + ((wh[k] >> (8 * c)) & 0xffu) * step;
        if (!(dl + step <= (double)lo[k][c]) || !(dh - step >= (double)hi[k][c])) ++qbad;
      }
    }
  }
  // The device's slab test on the decoded boxes never culls a child that
  // an exact test (double precision, axis-parallel components handled as
  // "origin inside the slab") keeps, over random rays: a quarter of them
  // axis-parallel with one or two zero components, some with the origin on
  // a vertex's plane. The device takes the reciprocal +-2^64 for components
  // below 2^-64 (slab_rcp).
  srand(7);
  const size_t nn = b.nodes.size() < 4000 ? b.nodes.size() : 4000;
  for (int r = 0; r < 600; ++r) {
    float o[3], d[3], inv[3], oi[3];
    for (int k = 0; k < 3; ++k) {
      o[k] = bmin[k] + (bmax[k] - bmin[k]) * (1.4f * rand() / (float)RAND_MAX - 0.2f);
      d[k] = 2.0f * rand() / (float)RAND_MAX - 1.0f;
    }
    if (r % 4 == 0) {
      const int k0 = (r / 4) % 3;
      d[k0] = (r % 12 == 0) ? -0.0f : 0.0f;
      if (r % 8 == 0) d[(k0 + 1) % 3] = 0.0f;
      if (r % 16 == 0) o[k0] = pos[9 * (rand() % n) + k0];
    }
    for (int k = 0; k < 3; ++k) {
      inv[k] = std::fabs(d[k]) < 0x1p-64f ? std::copysign(0x1p64f, d[k]) : 1.0f / d[k];
      oi[k] = o[k] * inv[k];
    }
    for (size_t i = 0; i < nn; ++i) {
      const oxr::Node4& nd = b.nodes[i];
      const oxr::Node4Q& nq = q[i];
      const float* lo[3] = {nd.lo_x, nd.lo_y, nd.lo_z};
      const float* hi[3] = {nd.hi_x, nd.hi_y, nd.hi_z};
      const uint32_t wl[3] = {nq.qlo_x, nq.qlo_y, nq.qlo_z}, wh[3] = {nq.qhi_x, nq.qhi_y, nq.qhi_z};
      float sk[3], ok[3];
      for (int k = 0; k < 3; ++k) {
        sk[k] = std::ldexp(inv[k], (int)(int8_t)((nq.exps >> (8 * k)) & 0xffu));
        ok[k] = std::fma(nq.org[k], inv[k], -oi[k]);
      }
      for (int c = 0; c < 4; ++c) {
        if (nd.child[c] == oxr::NODE_EMPTY) continue;
        double n0 = 0.0, f0 = 1e30;
        bool inside = true;
        float n1 = 0.0f, f1 = 1e30f;
        for (int k = 0; k < 3; ++k) {
          if (d[k] == 0.0f) {
            if (o[k] < lo[k][c] || o[k] > hi[k][c]) inside = false;
          } else {
            double a = ((double)lo[k][c] - o[k]) / d[k], bb = ((double)hi[k][c] - o[k]) / d[k];
            n0 = std::fmax(n0, std::fmin(a, bb));
            f0 = std::fmin(f0, std::fmax(a, bb));
          }
          float qa = std::fma((float)((wl[k] >> (8 * c)) & 0xffu), sk[k], ok[k]);
          float qb = std::fma((float)((wh[k] >> (8 * c)) & 0xffu), sk[k], ok[k]);
          bool neg = std::signbit(inv[k]);
          n1 = std::fmax(n1, neg ? qb : qa); f1 = std::fmin(f1, neg ? qa : qb);
        }
        const float eps = 1.0f + 1.0f / 65536.0f;
        if (inside && n0 <= f0 && !(n1 <= f1 * eps)) ++culled;
      }
    }
  }
  printf("%zu %zu %u %ld %ld %ld %ld\n", b.nodes.size(), b.tris.size(), b.max_depth, bad, inleaf, qbad, culled);
  FILE* g = fopen(argv[3], "wb");
  for (const auto& t : b.tris) fwrite(&t.gid, 4, 1, g);
  fclose(g);
  return 0;
}
'''


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("bvh")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", f"-I{CSRC}", str(src), os.path.join(CSRC, "bvh.cpp"),
                    "-o", str(exe)], check=True)
    return d, exe


def _soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3), dtype=np.float32)
    off = (rng.random((n, 6), dtype=np.float32) - 0.5) * 0.01
    return np.concatenate([c, c + off[:, :3], c + off[:, 3:]], 1).astype(np.float32)


@pytest.mark.parametrize("n", [1000, 150_000, 260_000, 1_200_000])
def test_every_triangle_in_one_leaf_any_thread_count(harness, n):
    d, exe = harness
    pos = _soup(n, n)
    pf = d / f"p{n}.f32"
    pos.tofile(pf)
    outs = []
    for threads in (1, 8):
        gf = d / f"g{n}_{threads}.u32"
        r = subprocess.run([str(exe), str(pf), str(threads), str(gf)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        nodes, ntris, depth, bad, inleaf, qbad, culled = map(int, r.stdout.split())
        gids = np.fromfile(gf, dtype=np.uint32)
        assert ntris == n and inleaf == n and bad == 0
        assert qbad == 0 and culled == 0
        assert np.array_equal(np.sort(gids), np.arange(n, dtype=np.uint32))
        outs.append((nodes, depth, gids))
    assert outs[0][1] < 40
    # Same tree for any thread count (leaf order included).
    assert outs[0][0] == outs[1][0] and np.array_equal(outs[0][2], outs[1][2])


@pytest.mark.parametrize("case", ["far_from_origin", "flat", "mixed_scale"])
def test_quantised_nodes_contain_their_boxes(harness, case):
    """The 64-byte device nodes: boxes decoded from the byte planes hold the
    exact boxes with a margin of one step, and the float slab test on them
    keeps every child the exact test keeps, also for small triangles far
    from the origin (the step floor of a few ulps), flat scenes (a zero
    extent axis) and nodes spanning tiny and large triangles."""
    d, exe = harness
    rng = np.random.default_rng(11)
    n = 20000
    if case == "far_from_origin":
        pos = _soup(n, 3) * 1e-3 + np.float32(1000.0)
    elif case == "flat":
        pos = _soup(n, 4)
        pos[:, 2::3] = 0.25
    else:
        pos = _soup(n, 5)
        big = rng.random(n) < 0.01
        pos[big, 3:] = pos[big, :3].repeat(2, 0).reshape(-1, 6) + rng.normal(size=(big.sum(), 6)).astype(np.float32)
    pos = np.ascontiguousarray(pos, np.float32)
    pf = d / f"q_{case}.f32"
    pos.tofile(pf)
    gf = d / f"qg_{case}.u32"
    r = subprocess.run([str(exe), str(pf), "8", str(gf)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    nodes, ntris, depth, bad, inleaf, qbad, culled = map(int, r.stdout.split())
    assert ntris == n and bad == 0
    assert qbad == 0
    assert culled == 0



@pytest.mark.parametrize("n", [5, 1000, 260_000])
def test_linear_bvh_every_triangle_in_one_leaf(harness, n):
    """The fast (linear, Morton-order) build: every triangle in exactly one
    leaf, every leaf box holds its triangles, the quantised nodes hold
    their boxes; on a soup with many equal codes too (a dense clump)."""
    d, exe = harness
    pos = _soup(n, n + 1)
    pos[: n // 3] = pos[: n // 3] * 1e-5 + np.float32(0.5)  # a clump: equal Morton codes, median splits
    pf = d / f"l{n}.f32"
    pos.tofile(pf)
    gf = d / f"lg{n}.u32"
    r = subprocess.run([str(exe), str(pf), "8", str(gf), "fast"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    nodes, ntris, depth, bad, inleaf, qbad, culled = map(int, r.stdout.split())
    gids = np.fromfile(gf, dtype=np.uint32)
    assert ntris == n and inleaf == n and bad == 0
    assert qbad == 0 and culled == 0
    assert np.array_equal(np.sort(gids), np.arange(n, dtype=np.uint32))
    assert depth < 40


DUMP = r'''
#include "bvh.h"
#include <cstdio>
#include <vector>
int main(int argc, char** argv) {
  FILE* f = fopen(argv[1], "rb");
  fseek(f, 0, SEEK_END); long sz = ftell(f); fseek(f, 0, SEEK_SET);
  std::vector<float> pos(sz / 4);
  if (fread(pos.data(), 1, sz, f) != (size_t)sz) return 2;
  fclose(f);
  uint32_t n = pos.size() / 9;
  std::vector<uint32_t> flags(n, 0); std::vector<float> alpha(n, 1.0f);
  oxr::Bvh4 b = oxr::build_bvh4(pos.data(), n, flags.data(), alpha.data(), 4, 8, false);
  std::vector<oxr::Node4Q> q = oxr::quantize_nodes(b.nodes);
  FILE* o = fopen(argv[2], "wb"); fwrite(q.data(), sizeof(oxr::Node4Q), q.size(), o); fclose(o);
  o = fopen(argv[3], "wb"); fwrite(b.tris.data(), sizeof(oxr::TriRec), b.tris.size(), o); fclose(o);
  return 0;
}
'''


def test_axis_parallel_rays_walk_as_few_nodes_as_their_neighbours(tmp_path, oracle):
    """Rays with an exactly zero direction component (flat ground, mirror
    bounces) on the device tree of a 2 x 192^2-triangle heightfield, walked
    in the device's closest-hit and occlusion orders (oxrcpu_bvh4_count /
    _any, whose slab reciprocal is the device's): the same nodes and
    triangles as the rays with that component set to +-1e-9 (a direction
    change of 1e-9 radians), and a small mean. An infinite reciprocal makes
    the node's slab offset NaN and the axis cull nothing: such a ray then
    fetches every node its other two slabs overlap (hundreds to thousands
    here)."""
    src = tmp_path / "d.cpp"
    src.write_text(DUMP)
    exe = tmp_path / "d"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", f"-I{CSRC}", str(src), os.path.join(CSRC, "bvh.cpp"),
                    "-o", str(exe)], check=True)
    import oxr.scenes as S
    g = 192
    xs = np.linspace(0.0, 1.0, g + 1, dtype=np.float32)
    X, Z = np.meshgrid(xs, xs, indexing="xy")
    Y = (0.1 * S.fbm(X, Z, 3)).astype(np.float32)
    P = np.stack([X, Y, Z], -1)
    a, b_, c, d = P[:-1, :-1], P[:-1, 1:], P[1:, :-1], P[1:, 1:]
    pos = np.concatenate([np.stack([a, c, b_], 2).reshape(-1, 9), np.stack([b_, c, d], 2).reshape(-1, 9)])
    pos = np.ascontiguousarray(pos, np.float32)
    pf, nf, tf = tmp_path / "p.f32", tmp_path / "n.bin", tmp_path / "t.bin"
    pos.tofile(pf)
    subprocess.run([str(exe), str(pf), str(nf), str(tf)], check=True, timeout=120)
    nodes = np.fromfile(nf, np.float32)
    tris = np.fromfile(tf, np.float32)
    rng = np.random.default_rng(5)
    m = 3000
    o = np.stack([rng.uniform(0.05, 0.95, m), rng.uniform(0.0, 0.3, m), rng.uniform(0.05, 0.95, m)], 1)
    k0 = np.arange(m) % 3
    d = rng.normal(size=(m, 3))
    d[np.arange(m), k0] = 0.0
    sel = np.arange(0, m, 7)
    d[sel, (k0[sel] + 1) % 3] = 0.0  # two zero components
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((m, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    near = rays.copy()
    near[np.arange(m), 4 + k0] = np.where(np.arange(m) % 2 == 0, 1e-9, -1e-9)
    near[sel, 4 + (k0[sel] + 1) % 3] = 1e-9
    shadow = rays.copy()
    shadow[:, 3] = np.frombuffer(np.uint32(0xFFFFFFFF).tobytes(), np.float32)[0]
    shadow[:, 7] = 0.5
    shadow_near = near.copy()
    shadow_near[:, 3] = shadow[:, 3]
    shadow_near[:, 7] = 0.5
    for walk, r0, r1 in ((oracle.bvh4_count, rays, near), (oracle.bvh4_count_any, shadow, shadow_near)):
        c0 = walk(nodes, tris, r0).astype(np.int64)
        c1 = walk(nodes, tris, r1).astype(np.int64)
        assert np.abs(c0 - c1).max() <= 4, np.abs(c0 - c1).max()
        assert c0[:, 0].mean() < 80 and c0[:, 0].max() < 400, (c0[:, 0].mean(), c0[:, 0].max())
