"""Host-side BVH builder (oxideav-render_amd/csrc/bvh.cpp), compiled with a
small harness: every triangle lands in exactly one leaf, every leaf box
holds its triangles, and the tree is the same with 1 and 8 threads —
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
  std::vector<uint32_t> flags(n, 0); std::vector<float> alpha(n, 1.0f);
  oxr::Bvh4 b = oxr::build_bvh4(pos.data(), n, flags.data(), alpha.data(), 4, atoi(argv[2]));
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
  printf("%zu %zu %u %ld %ld\n", b.nodes.size(), b.tris.size(), b.max_depth, bad, inleaf);
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
        nodes, ntris, depth, bad, inleaf = map(int, r.stdout.split())
        gids = np.fromfile(gf, dtype=np.uint32)
        assert ntris == n and inleaf == n and bad == 0
        assert np.array_equal(np.sort(gids), np.arange(n, dtype=np.uint32))
        outs.append((nodes, depth, gids))
    assert outs[0][1] < 40
    # Same tree for any thread count (leaf order included).
    assert outs[0][0] == outs[1][0] and np.array_equal(outs[0][2], outs[1][2])
