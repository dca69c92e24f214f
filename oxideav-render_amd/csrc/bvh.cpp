// Binned-SAH BVH build and 4-wide collapse (see bvh.h).
#include "bvh.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <limits>
#include <memory>
#include <mutex>
#include <thread>

namespace oxr {
namespace {

constexpr int NBINS = 32;
// SAH cost of a split relative to one triangle test. The binary tree is
// collapsed to 4-wide nodes whose hit leaves a wave tests in one divergent
// loop, while a visit's box tests run on every lane. Measured over 0.1 to 3
// (one MI355X, path-traced frames): 0.1-0.3 were fastest, against 1 the
// terrain +8% (triangle tests per ray 8.2 -> 5.1, node visits 37 -> 39),
// the material sweep +1%, the glossy caustic +1.5%; 2 and 3 were 8-26%
// slower.
constexpr double SAH_TRAVERSAL = 0.2;
constexpr float INF = std::numeric_limits<float>::infinity();

struct Box {
  float lo[3] = {INF, INF, INF};
  float hi[3] = {-INF, -INF, -INF};
  void grow(const Box& b) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::fmin(lo[k], b.lo[k]);
      hi[k] = std::fmax(hi[k], b.hi[k]);
    }
  }
  void grow_pt(const float* p) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::fmin(lo[k], p[k]);
      hi[k] = std::fmax(hi[k], p[k]);
    }
  }
  double area() const {
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = hi[k] >= lo[k] ? (double)hi[k] - (double)lo[k] : 0.0;
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
  }
  bool valid() const { return hi[0] >= lo[0]; }
};

struct Item {
  Box b;
  float c[3];
  uint32_t g;
};

struct BNode {
  Box b;
  uint32_t left = 0, right = 0;  // inner: child node ids
  uint32_t first = 0, count = 0;  // leaf: item range (count > 0)
};

struct Builder {
  std::vector<Item>& items;
  std::vector<BNode>& nodes;
  std::atomic<uint32_t> next_node{1};
  int max_leaf;
  // Work queue of subtrees for the worker threads.
  struct Task {
    uint32_t node, lo, hi;
  };
  std::mutex mu;
  std::condition_variable cv;
  std::vector<Task> queue;
  int busy = 0;
  bool done = false;

  Builder(std::vector<Item>& it, std::vector<BNode>& nd, int ml, int threads)
      : items(it), nodes(nd), max_leaf(ml), par(threads) {}

  // Nodes of at least BIG items (the top levels, which would otherwise run
  // on one thread) split their passes over `par` threads; the item order
  // they produce does not depend on the thread count.
  static constexpr uint32_t BIG = 1u << 20;
  int par = 1;
  std::vector<Item> tmp;  // scatter target of the big-node partitions

  template <class F>
  void par_for(uint32_t lo, uint32_t hi, F fn) {
    int T = std::max(1, par);
    std::vector<std::thread> ts;
    for (int t = 0; t < T; ++t) {
      uint32_t a = lo + (uint32_t)((uint64_t)(hi - lo) * t / T), b = lo + (uint32_t)((uint64_t)(hi - lo) * (t + 1) / T);
      if (T == 1) {
        fn(t, a, b);
        return;
      }
      ts.emplace_back([=, &fn] { fn(t, a, b); });
    }
    for (auto& th : ts) th.join();
  }

  void bounds(uint32_t lo, uint32_t hi, Box& b, Box& cb) {
    if (hi - lo < BIG) {
      for (uint32_t i = lo; i < hi; ++i) {
        b.grow(items[i].b);
        cb.grow_pt(items[i].c);
      }
      return;
    }
    std::vector<Box> pb(par), pc(par);
    par_for(lo, hi, [&](int t, uint32_t a, uint32_t e) {
      for (uint32_t i = a; i < e; ++i) {
        pb[t].grow(items[i].b);
        pc[t].grow_pt(items[i].c);
      }
    });
    for (int t = 0; t < par; ++t) {
      b.grow(pb[t]);
      cb.grow(pc[t]);
    }
  }

  // Per-axis bins of items [lo, hi): boxes and counts.
  void bin(uint32_t lo, uint32_t hi, const Box& cb, int nb, Box bb[3][NBINS], uint32_t bn[3][NBINS]) {
    auto pass = [&](uint32_t a, uint32_t e, Box lb[3][NBINS], uint32_t ln[3][NBINS]) {
      for (int axis = 0; axis < 3; ++axis) {
        float ext = cb.hi[axis] - cb.lo[axis];
        if (!(ext > 0.0f) || !std::isfinite(ext)) continue;
        float k = (float)nb / ext;
        for (uint32_t i = a; i < e; ++i) {
          int b = (int)((items[i].c[axis] - cb.lo[axis]) * k);
          b = std::min(std::max(b, 0), nb - 1);
          lb[axis][b].grow(items[i].b);
          ln[axis][b]++;
        }
      }
    };
    if (hi - lo < BIG) {
      pass(lo, hi, bb, bn);
      return;
    }
    struct Part {
      Box b[3][NBINS];
      uint32_t n[3][NBINS] = {};
    };
    std::vector<Part> parts(par);
    par_for(lo, hi, [&](int t, uint32_t a, uint32_t e) { pass(a, e, parts[t].b, parts[t].n); });
    for (int t = 0; t < par; ++t)
      for (int axis = 0; axis < 3; ++axis)
        for (int i = 0; i < nb; ++i) {
          bb[axis][i].grow(parts[t].b[axis][i]);
          bn[axis][i] += parts[t].n[axis][i];
        }
  }

  // Stable partition of a big node by the predicate, through `tmp`.
  template <class Pred>
  uint32_t big_partition(uint32_t lo, uint32_t hi, Pred left) {
    std::vector<uint32_t> nl(par, 0), nr(par, 0);
    par_for(lo, hi, [&](int t, uint32_t a, uint32_t e) {
      for (uint32_t i = a; i < e; ++i) (left(items[i]) ? nl[t] : nr[t])++;
    });
    uint32_t total_l = 0;
    for (int t = 0; t < par; ++t) total_l += nl[t];
    std::vector<uint32_t> ol(par), orr(par);
    uint32_t accl = lo, accr = lo + total_l;
    for (int t = 0; t < par; ++t) {
      ol[t] = accl;
      orr[t] = accr;
      accl += nl[t];
      accr += nr[t];
    }
    par_for(lo, hi, [&](int t, uint32_t a, uint32_t e) {
      uint32_t pl = ol[t], pr = orr[t];
      for (uint32_t i = a; i < e; ++i) tmp[left(items[i]) ? pl++ : pr++] = items[i];
    });
    par_for(lo, hi, [&](int, uint32_t a, uint32_t e) { std::copy(tmp.begin() + a, tmp.begin() + e, items.begin() + a); });
    return lo + total_l;
  }

  // Returns true and (axis, bin) of the best SAH split, or false for a leaf.
  bool find_split(uint32_t lo, uint32_t hi, const Box& nodes_bounds, const Box& cb, int& nb_out, int& axis_out,
                  int& bin_out) {
    uint32_t cnt = hi - lo;
    double best = std::numeric_limits<double>::infinity();
    int best_axis = -1, best_bin = -1;
    const Box& bounds = nodes_bounds;
    double parent_area = bounds.area();
    // Bins: as many as items up to NBINS (more bins than items only costs).
    const int nb = (int)std::min<uint32_t>(NBINS, std::max<uint32_t>(4u, cnt));
    // Big nodes: all three axes binned in one parallel pass.
    std::unique_ptr<Box[]> big_b;
    std::unique_ptr<uint32_t[]> big_n;
    if (cnt >= BIG) {
      big_b.reset(new Box[3 * NBINS]);
      big_n.reset(new uint32_t[3 * NBINS]());
      bin(lo, hi, cb, nb, reinterpret_cast<Box(*)[NBINS]>(big_b.get()),
          reinterpret_cast<uint32_t(*)[NBINS]>(big_n.get()));
    }
    for (int axis = 0; axis < 3; ++axis) {
      float ext = cb.hi[axis] - cb.lo[axis];
      if (!(ext > 0.0f) || !std::isfinite(ext)) continue;
      Box bb[NBINS];
      uint32_t bn[NBINS] = {0};
      if (cnt >= BIG) {
        for (int i = 0; i < nb; ++i) {
          bb[i] = big_b[axis * NBINS + i];
          bn[i] = big_n[axis * NBINS + i];
        }
      } else {
        float k = (float)nb / ext;
        for (uint32_t i = lo; i < hi; ++i) {
          int b = (int)((items[i].c[axis] - cb.lo[axis]) * k);
          b = std::min(std::max(b, 0), nb - 1);
          bb[b].grow(items[i].b);
          bn[b]++;
        }
      }
      double la[NBINS], ra[NBINS];
      uint32_t ln[NBINS], rn[NBINS];
      Box acc;
      uint32_t n = 0;
      for (int i = 0; i < nb; ++i) {
        acc.grow(bb[i]);
        n += bn[i];
        la[i] = acc.area();
        ln[i] = n;
      }
      acc = Box();
      n = 0;
      for (int i = nb - 1; i >= 0; --i) {
        acc.grow(bb[i]);
        n += bn[i];
        ra[i] = acc.area();
        rn[i] = n;
      }
      for (int i = 0; i < nb - 1; ++i) {
        if (ln[i] == 0 || rn[i + 1] == 0) continue;
        double c = la[i] * ln[i] + ra[i + 1] * rn[i + 1];
        if (c < best) {
          best = c;
          best_axis = axis;
          best_bin = i + 1;
        }
      }
    }
    // SAH leaf test: a triangle costs 1, a split SAH_TRAVERSAL.
    double leaf_cost = (double)cnt;
    double split_cost = parent_area > 0.0 ? SAH_TRAVERSAL + best / parent_area : SAH_TRAVERSAL + (double)cnt;
    if (cnt <= (uint32_t)max_leaf && (best_axis < 0 || leaf_cost <= split_cost)) return false;
    axis_out = best_axis;
    bin_out = best_bin;
    nb_out = nb;
    return true;
  }

  // Builds the subtree of `node` over items [lo, hi); pushes big
  // subtrees to the shared queue and recurses on small ones.
  void build(uint32_t node, uint32_t lo, uint32_t hi) {
    for (;;) {
      BNode& nd = nodes[node];
      Box b, cb;
      bounds(lo, hi, b, cb);
      nd.b = b;
      uint32_t cnt = hi - lo;
      int axis = -1, bin = -1, nbins = NBINS;
      bool split = cnt > 1 && find_split(lo, hi, b, cb, nbins, axis, bin);
      uint32_t mid = lo;
      if (split && axis >= 0) {
        float ext = cb.hi[axis] - cb.lo[axis];
        float k = (float)nbins / ext;
        float clo = cb.lo[axis];
        auto left = [&](const Item& x) {
          int bb = (int)((x.c[axis] - clo) * k);
          bb = std::min(std::max(bb, 0), nbins - 1);
          return bb < bin;
        };
        if (cnt >= BIG) {
          mid = big_partition(lo, hi, left);
        } else {
          auto it = std::partition(items.begin() + lo, items.begin() + hi, left);
          mid = (uint32_t)(it - items.begin());
        }
      }
      if (split && (axis < 0 || mid == lo || mid == hi)) {
        if (cnt <= (uint32_t)max_leaf) {
          split = false;
        } else {
          // Degenerate centroids: median split along the widest axis.
          int ax = 0;
          for (int kk = 1; kk < 3; ++kk)
            if (cb.hi[kk] - cb.lo[kk] > cb.hi[ax] - cb.lo[ax]) ax = kk;
          mid = lo + cnt / 2;
          std::nth_element(items.begin() + lo, items.begin() + mid, items.begin() + hi,
                           [&](const Item& a, const Item& c) { return a.c[ax] < c.c[ax] || (a.c[ax] == c.c[ax] && a.g < c.g); });
        }
      }
      if (!split) {
        if (cnt > (uint32_t)max_leaf) {
          // Can only happen for identical centroids with cnt > max_leaf:
          // split in the middle of the range.
          mid = lo + cnt / 2;
        } else {
          nd.first = lo;
          nd.count = cnt;
          return;
        }
      }
      uint32_t l = next_node.fetch_add(2);
      nd.left = l;
      nd.right = l + 1;
      nd.count = 0;
      // Hand the larger half to the pool when it is big, keep the other.
      uint32_t nl = mid - lo, nr = hi - mid;
      if (nl > nr) {
        if (nl > 65536) {
          push({l, lo, mid});
        } else {
          build(l, lo, mid);
        }
        node = l + 1;
        lo = mid;
      } else {
        if (nr > 65536) {
          push({l + 1, mid, hi});
        } else {
          build(l + 1, mid, hi);
        }
        node = l;
        hi = mid;
      }
    }
  }

  void push(Task t) {
    {
      std::lock_guard<std::mutex> g(mu);
      queue.push_back(t);
    }
    cv.notify_one();
  }

  void worker() {
    for (;;) {
      Task t;
      {
        std::unique_lock<std::mutex> g(mu);
        cv.wait(g, [&] { return !queue.empty() || (busy == 0 && queue.empty()) || done; });
        if (queue.empty()) {
          done = true;
          cv.notify_all();
          return;
        }
        t = queue.back();
        queue.pop_back();
        ++busy;
      }
      build(t.node, t.lo, t.hi);
      {
        std::lock_guard<std::mutex> g(mu);
        --busy;
      }
      cv.notify_all();
    }
  }
};

struct Collapser {
  const std::vector<BNode>& bn;
  const std::vector<Item>& items;
  const float* pos;
  const uint32_t* flags;
  const float* alpha;
  float pad;
  Bvh4& out;

  double area(uint32_t id) const { return bn[id].b.area(); }

  // Writes the 4-wide node for binary node `id` (an inner node) at out
  // index `slot`; returns depth of the subtree.
  uint32_t emit(uint32_t id, uint32_t slot) {
    uint32_t kids[4];
    int nk = 0;
    kids[nk++] = bn[id].left;
    kids[nk++] = bn[id].right;
    while (nk < 4) {
      int best = -1;
      double ba = -1.0;
      for (int i = 0; i < nk; ++i) {
        const BNode& c = bn[kids[i]];
        if (c.count == 0 && area(kids[i]) > ba) {
          ba = area(kids[i]);
          best = i;
        }
      }
      if (best < 0) break;
      uint32_t c = kids[best];
      kids[best] = bn[c].left;
      kids[nk++] = bn[c].right;
    }
    Node4 n4;
    for (int i = 0; i < 4; ++i) {
      n4.lo_x[i] = n4.lo_y[i] = n4.lo_z[i] = INF;
      n4.hi_x[i] = n4.hi_y[i] = n4.hi_z[i] = -INF;
      n4.child[i] = NODE_EMPTY;
      n4.count[i] = 0;
    }
    out.nodes[slot] = n4;
    uint32_t depth = 1;
    for (int i = 0; i < nk; ++i) {
      const BNode& c = bn[kids[i]];
      Node4& w = out.nodes[slot];
      w.lo_x[i] = c.b.lo[0] - pad;
      w.lo_y[i] = c.b.lo[1] - pad;
      w.lo_z[i] = c.b.lo[2] - pad;
      w.hi_x[i] = c.b.hi[0] + pad;
      w.hi_y[i] = c.b.hi[1] + pad;
      w.hi_z[i] = c.b.hi[2] + pad;
      if (c.count > 0) {
        w.child[i] = (uint32_t)out.tris.size();
        w.count[i] = c.count;
        for (uint32_t k = 0; k < c.count; ++k) out.tris.push_back(make_tri(items[c.first + k].g));
        out.num_leaves++;
        out.sah_cost += c.b.area() * (double)c.count;
      } else {
        uint32_t child_slot = (uint32_t)out.nodes.size();
        out.nodes.emplace_back();
        out.nodes[slot].child[i] = child_slot;
        out.nodes[slot].count[i] = 0;
        out.sah_cost += c.b.area();
        uint32_t d = emit(kids[i], child_slot);
        depth = std::max(depth, d + 1);
      }
    }
    return depth;
  }

  TriRec make_tri(uint32_t g) const {
    TriRec t;
    const float* p = pos + 9ull * g;
    for (int k = 0; k < 3; ++k) {
      t.v0[k] = p[k];
      t.v1[k] = p[3 + k];
      t.v2[k] = p[6 + k];
    }
    t.gid = g;
    t.flags = flags ? flags[g] : 0u;
    t.alpha = alpha ? alpha[g] : 1.0f;
    return t;
  }
};

// Linear BVH: items in Morton order of their centroids (10 bits per
// axis), split top-down where the highest differing code bit flips
// (Lauterbach et al., "Fast BVH Construction on GPUs", 2009; the radix-tree
// split of Karras, HPG 2012); equal codes split at the middle. Boxes are
// the unions of their items'. Linear in n after the sort: several times
// faster than the binned SAH build, with a looser tree.
struct Lbvh {
  std::vector<Item>& items;
  std::vector<BNode>& nodes;
  std::vector<uint32_t> code;
  uint32_t used = 1;
  int max_leaf;

  static uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | v << 16) & 0x030000ffu;
    v = (v | v << 8) & 0x0300f00fu;
    v = (v | v << 4) & 0x030c30c3u;
    v = (v | v << 2) & 0x09249249u;
    return v;
  }

  void sort() {
    const uint32_t n = (uint32_t)items.size();
    Box cb;
    for (const Item& it : items) cb.grow_pt(it.c);
    float s[3];
    for (int k = 0; k < 3; ++k) {
      float e = cb.hi[k] - cb.lo[k];
      s[k] = (std::isfinite(e) && e > 0.0f) ? 1023.0f / e : 0.0f;
    }
    // (code, index) pairs, LSD radix sorted by two 16-bit digits: stable,
    // so equal codes keep the items' input order.
    std::vector<uint64_t> a(n), t(n);
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t q[3];
      for (int k = 0; k < 3; ++k) {
        float v = (items[i].c[k] - cb.lo[k]) * s[k];
        q[k] = (std::isfinite(v) && v > 0.0f) ? (uint32_t)std::fmin(v, 1023.0f) : 0u;
      }
      const uint32_t key = spread10(q[0]) << 2 | spread10(q[1]) << 1 | spread10(q[2]);
      a[i] = (uint64_t)key << 32 | i;
    }
    std::vector<uint32_t> cnt(65537);
    for (int pass = 0; pass < 2; ++pass) {
      const int sh = 32 + 16 * pass;
      std::fill(cnt.begin(), cnt.end(), 0u);
      for (uint32_t i = 0; i < n; ++i) cnt[((a[i] >> sh) & 0xffff) + 1]++;
      if (cnt[1] == n) continue;  // every code has this digit 0
      for (int d = 0; d < 65536; ++d) cnt[d + 1] += cnt[d];
      for (uint32_t i = 0; i < n; ++i) t[cnt[(a[i] >> sh) & 0xffff]++] = a[i];
      a.swap(t);
    }
    std::vector<Item> sorted(n);
    code.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
      sorted[i] = items[(uint32_t)a[i]];
      code[i] = (uint32_t)(a[i] >> 32);
    }
    items.swap(sorted);
  }

  // Node `id` over items [lo, hi); returns its box.
  Box build(uint32_t id, uint32_t lo, uint32_t hi) {
    BNode& nd = nodes[id];
    if (hi - lo <= (uint32_t)max_leaf) {
      Box b;
      for (uint32_t i = lo; i < hi; ++i) b.grow(items[i].b);
      nd.b = b;
      nd.first = lo;
      nd.count = hi - lo;
      return b;
    }
    uint32_t mid;
    const uint32_t a = code[lo], z = code[hi - 1];
    if (a == z) {
      mid = lo + (hi - lo) / 2;
    } else {
      // First index whose code has the highest differing bit set.
      const int bit = 31 - __builtin_clz(a ^ z);
      const uint32_t m = 1u << bit;
      uint32_t l = lo, h = hi - 1;  // code[l] has the bit clear, code[h] set
      while (h - l > 1) {
        uint32_t c = l + (h - l) / 2;
        if (code[c] & m)
          h = c;
        else
          l = c;
      }
      mid = h;
    }
    const uint32_t l = used;
    used += 2;
    nd.left = l;
    nd.right = l + 1;
    nd.count = 0;
    Box b = build(l, lo, mid);
    b.grow(build(l + 1, mid, hi));
    nodes[id].b = b;
    return b;
  }
};

}  // namespace

Bvh4 build_bvh4(const float* pos, uint32_t n, const uint32_t* flags, const float* alpha, int max_leaf, int threads,
                bool fast) {
  auto t0 = std::chrono::steady_clock::now();
  Bvh4 out;
  if (n == 0) return out;
  max_leaf = std::min(std::max(max_leaf, 1), 8);
  std::vector<Item> items(n);
  float gmax = 0.0f;
  for (uint32_t g = 0; g < n; ++g) {
    Item& it = items[g];
    const float* p = pos + 9ull * g;
    for (int k = 0; k < 3; ++k) {
      it.b.lo[k] = std::fmin(p[k], std::fmin(p[3 + k], p[6 + k]));
      it.b.hi[k] = std::fmax(p[k], std::fmax(p[3 + k], p[6 + k]));
      it.c[k] = 0.5f * (it.b.lo[k] + it.b.hi[k]);
      if (std::isfinite(it.b.lo[k])) gmax = std::fmax(gmax, std::fabs(it.b.lo[k]));
      if (std::isfinite(it.b.hi[k])) gmax = std::fmax(gmax, std::fabs(it.b.hi[k]));
    }
    it.g = g;
  }
  // Absolute padding of every box: keeps the slab test conservative
  // against the rounding of the watertight triangle test.
  out.pad = std::ldexp(std::fmax(gmax, 1.0f), -20);
  std::vector<BNode> bn(2ull * n);
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = std::min(nt, 32);
  Builder b(items, bn, max_leaf, nt);
  if (fast) {
    Lbvh lb{items, bn, {}, 1u, max_leaf};
    lb.sort();
    lb.build(0, 0, n);
    b.next_node = lb.used;
  } else if (n >= Builder::BIG) {
    b.tmp.resize(n);
  }
  if (fast) {
    // built above
  } else if (n < 200000 || nt == 1) {
    // One thread: build() still hands subtrees of more than 65536 items to
    // the queue; drain it here.
    b.build(0, 0, n);
    while (!b.queue.empty()) {
      Builder::Task t = b.queue.back();
      b.queue.pop_back();
      b.build(t.node, t.lo, t.hi);
    }
  } else {
    b.push({0, 0, n});
    std::vector<std::thread> ts;
    for (int i = 0; i < nt; ++i) ts.emplace_back([&] { b.worker(); });
    for (auto& t : ts) t.join();
  }
  uint32_t used = b.next_node.load();
  bn.resize(used);
  out.nodes.reserve(used / 2 + 2);
  out.tris.reserve(n);
  Collapser c{bn, items, pos, flags, alpha, out.pad, out};
  out.nodes.emplace_back();
  if (bn[0].count > 0) {
    // The whole scene is one leaf: a root with a single leaf child.
    Node4 n4;
    for (int i = 0; i < 4; ++i) {
      n4.lo_x[i] = n4.lo_y[i] = n4.lo_z[i] = INF;
      n4.hi_x[i] = n4.hi_y[i] = n4.hi_z[i] = -INF;
      n4.child[i] = NODE_EMPTY;
      n4.count[i] = 0;
    }
    n4.lo_x[0] = bn[0].b.lo[0] - out.pad;
    n4.lo_y[0] = bn[0].b.lo[1] - out.pad;
    n4.lo_z[0] = bn[0].b.lo[2] - out.pad;
    n4.hi_x[0] = bn[0].b.hi[0] + out.pad;
    n4.hi_y[0] = bn[0].b.hi[1] + out.pad;
    n4.hi_z[0] = bn[0].b.hi[2] + out.pad;
    n4.child[0] = 0;
    n4.count[0] = bn[0].count;
    for (uint32_t k = 0; k < bn[0].count; ++k) out.tris.push_back(c.make_tri(items[bn[0].first + k].g));
    out.nodes[0] = n4;
    out.num_leaves = 1;
    out.max_depth = 1;
  } else {
    out.max_depth = c.emit(0, 0);
  }
  double root_area = bn[0].b.area();
  if (root_area > 0.0) out.sah_cost /= root_area;
  out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return out;
}

namespace {

// Largest f32 not above x (x finite).
float f32_down(double x) {
  float f = (float)x;
  if ((double)f > x) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
  return f;
}

// One axis of a node: org, the exponent and its children's byte planes.
struct AxisQ {
  float org = 0.0f;
  int e = 0;
  uint32_t lo = 0, hi = 0;
};

AxisQ quantize_axis(const float* lo, const float* hi, const uint32_t* child) {
  double L = std::numeric_limits<double>::infinity(), H = -L;
  bool unbounded = false;
  for (int c = 0; c < 4; ++c)
    if (child[c] != NODE_EMPTY) {
      unbounded = unbounded || !std::isfinite(lo[c]) || !std::isfinite(hi[c]);
      L = std::min(L, (double)lo[c]);
      H = std::max(H, (double)hi[c]);
    }
  AxisQ a;
  if (unbounded) {
    // A child box with a non-finite plane (a triangle with a non-finite
    // corner): a NaN corner makes every plane of this axis NaN, which the
    // slab test's min/max ignore, so no child is culled on this axis.
    a.org = std::numeric_limits<float>::quiet_NaN();
    for (int c = 0; c < 4; ++c)
      if (child[c] != NODE_EMPTY) a.hi |= 255u << (8 * c);
    return a;
  }
  if (!(L <= H)) return a;  // no child: never decoded
  // A step of at least a few ulps of the coordinates, and fine enough that
  // the node's extent spans at most ~250 of them.
  const double mag = std::max(std::fabs(L), std::fabs(H));
  int e = std::ilogb(std::max(mag, 1e-30)) - 21;
  while (std::ldexp(250.0, e) < H - L) ++e;
  e = std::max(e, -126);
  for (;; ++e) {
    const double step = std::ldexp(1.0, e);
    const float org = f32_down(L - step);
    // Widest plane: ceil((H - org) / step) + 1 must fit in a byte.
    if (std::ceil(((double)H - (double)org) / step) + 1.0 > 255.0) continue;
    a.org = org;
    a.e = e;
    for (int c = 0; c < 4; ++c) {
      if (child[c] == NODE_EMPTY) continue;
      const double ql = std::floor(((double)lo[c] - (double)org) / step) - 1.0;
      const double qh = std::ceil(((double)hi[c] - (double)org) / step) + 1.0;
      a.lo |= (uint32_t)std::max(ql, 0.0) << (8 * c);
      a.hi |= (uint32_t)std::min(qh, 255.0) << (8 * c);
    }
    return a;
  }
}

}  // namespace

std::vector<Node4Q> quantize_nodes(const std::vector<Node4>& nodes) {
  std::vector<Node4Q> out(nodes.size());
  for (size_t i = 0; i < nodes.size(); ++i) {
    const Node4& n = nodes[i];
    Node4Q& q = out[i];
    const AxisQ x = quantize_axis(n.lo_x, n.hi_x, n.child);
    const AxisQ y = quantize_axis(n.lo_y, n.hi_y, n.child);
    const AxisQ z = quantize_axis(n.lo_z, n.hi_z, n.child);
    q.org[0] = x.org;
    q.org[1] = y.org;
    q.org[2] = z.org;
    q.exps = ((uint32_t)(uint8_t)(int8_t)x.e) | ((uint32_t)(uint8_t)(int8_t)y.e << 8) |
             ((uint32_t)(uint8_t)(int8_t)z.e << 16);
    q.qlo_x = x.lo;
    q.qhi_x = x.hi;
    q.qlo_y = y.lo;
    q.qhi_y = y.hi;
    q.qlo_z = z.lo;
    q.qhi_z = z.hi;
    q.counts = 0;
    for (int c = 0; c < 4; ++c) q.counts |= n.count[c] << (8 * c);  // leaves hold at most 8 triangles
    q.pad = 0;
    for (int c = 0; c < 4; ++c) q.child[c] = n.child[c];
  }
  return out;
}

}  // namespace oxr
