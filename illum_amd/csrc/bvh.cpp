// Host BVH build: binned SAH over triangle bounding boxes, flattened to a
// threaded (stackless) layout in depth-first pre-order.  Replaces the Fast-BVH
// constructor used by the reference (illum.cpp:25, illum.cpp:333-338).
//
// Layout (consumed by the traversal kernels):
//   * inner node i: its first child is i + 1; its second child follows the
//     whole first subtree; miss(i) = the node after the whole subtree of i;
//   * leaf: triangles [first, first + count) of the reordered triangle array.
// Boxes are padded by 4e-6 of the scene's largest coordinate magnitude so a
// ray that the FP32 watertight predicate reports as hitting a triangle always
// passes through every ancestor box with a wide margin against rounding.
//
// A mesh of several bodies (contiguous facet ranges, HostGeometry::body_off)
// gets one tree per body, concatenated: body b's trees start at root_thr[b]
// (threaded) and root_wide[b] (two-wide), its threaded walk ends at miss link
// -1, and its leaves index its own facets.  A walk from a body's root sees
// only that body.
#include "internal.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <atomic>
#include <cstring>
#include <string>
#include <thread>

namespace illum {

namespace {

struct Box {
  float lo[3], hi[3];
  Box() { for (int c = 0; c < 3; ++c) { lo[c] = FLT_MAX; hi[c] = -FLT_MAX; } }
  void grow(const Box &b) {
    for (int c = 0; c < 3; ++c) { lo[c] = std::min(lo[c], b.lo[c]); hi[c] = std::max(hi[c], b.hi[c]); }
  }
  void grow(const float *p) {
    for (int c = 0; c < 3; ++c) { lo[c] = std::min(lo[c], p[c]); hi[c] = std::max(hi[c], p[c]); }
  }
  double area() const {
    double dx = (double)hi[0] - lo[0], dy = (double)hi[1] - lo[1], dz = (double)hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
  }
};

struct BuildNode {
  Box box;
  int left = -1, right = -1;  // children (build indices)
  int first = 0, count = 0;   // leaf range in the permutation
};

struct Builder {
  const HostGeometry &g;
  int leaf_size;
  std::vector<Box> tbox;
  std::vector<float> tcen;   // centroid of each triangle's box
  std::vector<int> perm;
  std::vector<int> scratch;  // the partition's right-hand side, reused
  std::vector<BuildNode> nodes;
  int max_depth = 0;

  Builder(const HostGeometry &g_, int ls) : g(g_), leaf_size(ls) {}

  static constexpr int NB = 32;  // SAH bins per axis

  // Node over perm[first, first + count), whose triangles' boxes span b and
  // whose box centroids span cb (the root computes both; a child takes them
  // from its parent's bins and partition pass, the same values, since min
  // and max are exact).
  //
  // Binned SAH over the centroid box, NB bins per axis: one pass bins all
  // three axes (per bin: count and triangle box); the split with the least
  // cost (left area x count + right area x count) is taken, ties to the
  // lower axis and bin.  A split just past an empty bin costs the same as
  // the split before it, which wins the tie, so only splits past non-empty
  // bins are evaluated.  The range is then partitioned stably.
  int build(int first, int count, int depth, const Box &b, const Box &cb) {
    max_depth = std::max(max_depth, depth);
    int id = (int)nodes.size();
    nodes.emplace_back();
    nodes[id].box = b;
    if (count <= leaf_size) {
      nodes[id].first = first;
      nodes[id].count = count;
      return id;
    }
    Box bins[3][NB];
    int cnt[3][NB] = {};
    double lo[3], scale[3];
    bool ok[3];
    for (int axis = 0; axis < 3; ++axis) {
      lo[axis] = cb.lo[axis];
      const double hi = cb.hi[axis];
      ok[axis] = hi > lo[axis];
      scale[axis] = ok[axis] ? NB / (hi - lo[axis]) : 0.0;
    }
    for (int k = first; k < first + count; ++k) {
      const int t = perm[k];
      const float *c = &tcen[3 * t];
      for (int axis = 0; axis < 3; ++axis) {
        if (!ok[axis]) continue;
        int bi = (int)((c[axis] - lo[axis]) * scale[axis]);
        bi = std::min(NB - 1, std::max(0, bi));
        cnt[axis][bi]++;
        bins[axis][bi].grow(tbox[t]);
      }
    }
    int best_axis = -1, best_split = -1;
    double best_cost = DBL_MAX;
    for (int axis = 0; axis < 3; ++axis) {
      if (!ok[axis]) continue;
      double rarea[NB];
      int rcnt[NB];
      Box acc;
      int ac = 0;
      for (int s = NB - 1; s > 0; --s) {
        if (cnt[axis][s]) {
          acc.grow(bins[axis][s]);
          ac += cnt[axis][s];
        }
        rcnt[s] = ac;
        // needed only past a non-empty bin s - 1 (see above)
        if (cnt[axis][s - 1]) rarea[s] = acc.area();
      }
      Box lacc;
      int lc = 0;
      for (int s = 1; s < NB; ++s) {
        if (!cnt[axis][s - 1]) continue;
        lacc.grow(bins[axis][s - 1]);
        lc += cnt[axis][s - 1];
        if (rcnt[s] == 0) continue;
        double cost = lacc.area() * lc + rarea[s] * rcnt[s];
        if (cost < best_cost) { best_cost = cost; best_axis = axis; best_split = s; }
      }
    }
    int mid = -1;
    Box lb, lcb, rb, rcb;
    if (best_axis >= 0) {
      const int ax = best_axis;
      // stable partition: left elements compacted in place, right ones
      // through the scratch buffer; the children's centroid boxes on the way
      int nl = 0, nr = 0;
      for (int k = first; k < first + count; ++k) {
        const int t = perm[k];
        const float *c = &tcen[3 * t];
        int bi = (int)((c[ax] - lo[ax]) * scale[ax]);
        bi = std::min(NB - 1, std::max(0, bi));
        if (bi < best_split) { perm[first + nl++] = t; lcb.grow(c); }
        else { scratch[nr++] = t; rcb.grow(c); }
      }
      std::copy(scratch.begin(), scratch.begin() + nr, perm.begin() + first + nl);
      mid = first + nl;
      for (int s = 0; s < best_split; ++s) if (cnt[ax][s]) lb.grow(bins[ax][s]);
      for (int s = best_split; s < NB; ++s) if (cnt[ax][s]) rb.grow(bins[ax][s]);
    }
    if (mid < 0 || mid == first || mid == first + count) {
      // all centroids coincide (or one side is empty): split by index
      mid = first + count / 2;
      lb = rb = lcb = rcb = Box();
      for (int k = first; k < mid; ++k) { lb.grow(tbox[perm[k]]); lcb.grow(&tcen[3 * perm[k]]); }
      for (int k = mid; k < first + count; ++k) { rb.grow(tbox[perm[k]]); rcb.grow(&tcen[3 * perm[k]]); }
    }
    int l = build(first, mid - first, depth + 1, lb, lcb);
    int r = build(mid, first + count - mid, depth + 1, rb, rcb);
    nodes[id].left = l;
    nodes[id].right = r;
    return id;
  }

  int build_root(int count) {
    scratch.resize(count);
    Box b, cb;
    for (int k = 0; k < count; ++k) {
      b.grow(tbox[perm[k]]);
      cb.grow(&tcen[3 * perm[k]]);
    }
    return build(0, count, 0, b, cb);
  }
};

inline float as_float(int32_t i) { float f; std::memcpy(&f, &i, 4); return f; }

}  // namespace

// One tree over facets [f0, f1); leaf ranges index the reordered facets of
// that range (0-based within it).
static HostBVH build_one(const HostGeometry &g, int64_t f0, int64_t f1, int leaf_size,
                         float pad) {
  const int64_t nf = f1 - f0;
  Builder B(g, leaf_size);
  B.tbox.resize(nf);
  B.tcen.resize(3 * nf);
  B.perm.resize(nf);
  for (int64_t k = 0; k < nf; ++k) {
    const int64_t t = f0 + k;
    Box b;
    for (int v = 0; v < 3; ++v) b.grow(&g.tri[9 * t + 3 * v]);
    B.tbox[k] = b;
    for (int c = 0; c < 3; ++c) B.tcen[3 * k + c] = 0.5f * (b.lo[c] + b.hi[c]);
    B.perm[k] = (int)k;
  }
  B.nodes.reserve(2 * nf / std::max(1, leaf_size) + 16);
  B.build_root((int)nf);

  // Flatten in DFS pre-order, computing miss links.
  HostBVH h;
  h.leaf_size = leaf_size;
  h.depth = B.max_depth;
  const int nn = (int)B.nodes.size();
  h.nnodes = nn;
  h.nodes.resize(8 * (size_t)nn);
  h.tris.resize(12 * (size_t)nf);
  std::vector<int> order;  // build index in pre-order
  order.reserve(nn);
  std::vector<int> flat_of(nn, -1);
  std::vector<int> stack = {0};
  while (!stack.empty()) {
    int id = stack.back();
    stack.pop_back();
    flat_of[id] = (int)order.size();
    order.push_back(id);
    if (B.nodes[id].left >= 0) {
      stack.push_back(B.nodes[id].right);
      stack.push_back(B.nodes[id].left);
    }
  }
  // subtree size -> miss link = pre-order index + subtree size
  std::vector<int> subtree(nn, 1);
  for (int k = nn - 1; k >= 0; --k) {
    int id = order[k];
    if (B.nodes[id].left >= 0) subtree[id] = 1 + subtree[B.nodes[id].left] + subtree[B.nodes[id].right];
  }
  for (int k = 0; k < nn; ++k) {
    const BuildNode &bn = B.nodes[order[k]];
    int miss = k + subtree[order[k]];
    if (miss >= nn) miss = -1;
    int32_t info;
    if (bn.left >= 0) {
      info = -1;
    } else {
      if (bn.count >= 128) throw Error(ILLUM_ERR_CAPACITY, "leaf too large");
      info = (bn.count << 24) | bn.first;
    }
    float *o = &h.nodes[8 * (size_t)k];
    o[0] = bn.box.lo[0] - pad; o[1] = bn.box.lo[1] - pad; o[2] = bn.box.lo[2] - pad;
    o[3] = as_float(miss);
    o[4] = bn.box.hi[0] + pad; o[5] = bn.box.hi[1] + pad; o[6] = bn.box.hi[2] + pad;
    o[7] = as_float(info);
  }
  // Two-wide layout: inner nodes in pre-order, each holding both child boxes.
  {
    std::vector<int> inner_of(nn, -1);
    int ninner = 0;
    for (int k = 0; k < nn; ++k)
      if (B.nodes[order[k]].left >= 0) inner_of[order[k]] = ninner++;
    auto leaf_info = [&](const BuildNode &bn) {
      if (bn.count >= 128) throw Error(ILLUM_ERR_CAPACITY, "leaf too large");
      return (int32_t)((bn.count << 24) | bn.first);
    };
    auto put_box = [&](float *o, const Box &b, int32_t ref) {
      o[0] = b.lo[0] - pad; o[1] = b.lo[1] - pad; o[2] = b.lo[2] - pad; o[3] = as_float(ref);
      o[4] = b.hi[0] + pad; o[5] = b.hi[1] + pad; o[6] = b.hi[2] + pad; o[7] = 0.0f;
    };
    if (ninner == 0) {  // root is a leaf: one inner root, empty right box
      h.ninner = 1;
      h.wide.assign(16, 0.0f);
      put_box(&h.wide[0], B.nodes[0].box, ~leaf_info(B.nodes[0]));
      Box empty;  // lo = +max, hi = -max: never entered
      for (int c = 0; c < 3; ++c) { h.wide[8 + c] = empty.lo[c]; h.wide[12 + c] = empty.hi[c]; }
      h.wide[11] = as_float(~leaf_info(B.nodes[0]));
    } else {
      h.ninner = ninner;
      h.wide.assign(16 * (size_t)ninner, 0.0f);
      for (int k = 0; k < nn; ++k) {
        const BuildNode &bn = B.nodes[order[k]];
        if (bn.left < 0) continue;
        float *o = &h.wide[16 * (size_t)inner_of[order[k]]];
        const int ch[2] = {bn.left, bn.right};
        for (int s = 0; s < 2; ++s) {
          const BuildNode &cn = B.nodes[ch[s]];
          const int32_t ref = cn.left >= 0 ? inner_of[ch[s]] : ~leaf_info(cn);
          put_box(o + 8 * s, cn.box, ref);
        }
      }
    }
  }
  // Four-wide layout: the binary tree collapsed so each node holds up to 4
  // children -- a binary node's two children, the inner child with the
  // largest box area replaced by its own two children until there are four.
  // Node = 8 float4: lo.x, lo.y, lo.z, hi.x, hi.y, hi.z of children 0..3
  // (SoA), child refs (>= 0 node, < 0 ~leaf info), (child count, 0, 0, 0).
  {
    // up to four children, in order (fixed records: no allocation per node)
    struct Kids {
      int c[4];
      int n;
      size_t size() const { return (size_t)n; }
      int operator[](size_t q) const { return c[q]; }
    };
    std::vector<Kids> kids;
    kids.reserve(nn / 2 + 1);
    std::vector<int> n4_of(nn, -1);
    auto expand = [&](int id) {
      Kids ch{{B.nodes[id].left, B.nodes[id].right, 0, 0}, 2};
      while (ch.n < 4) {
        int best = -1;
        double ba = -1.0;
        for (int q = 0; q < ch.n; ++q)
          if (B.nodes[ch.c[q]].left >= 0 && B.nodes[ch.c[q]].box.area() > ba) {
            ba = B.nodes[ch.c[q]].box.area();
            best = q;
          }
        if (best < 0) break;
        const int m = ch.c[best];
        for (int q = best; q + 1 < ch.n; ++q) ch.c[q] = ch.c[q + 1];  // erase, keeping order
        ch.n -= 1;
        ch.c[ch.n++] = B.nodes[m].left;
        ch.c[ch.n++] = B.nodes[m].right;
      }
      return ch;
    };
    int depth4 = 0;
    if (B.nodes[0].left < 0) {
      kids.push_back(Kids{{0, 0, 0, 0}, 1});  // root is a leaf
    } else {
      std::vector<std::pair<int, int>> st = {{0, 0}};
      while (!st.empty()) {
        const int id = st.back().first, dep = st.back().second;
        st.pop_back();
        depth4 = std::max(depth4, dep);
        n4_of[id] = (int)kids.size();
        kids.push_back(expand(id));
        const Kids ch = kids.back();
        for (int q = ch.n - 1; q >= 0; --q)
          if (B.nodes[ch.c[q]].left >= 0) st.push_back({ch.c[q], dep + 1});
      }
    }
    h.ninner4 = (int64_t)kids.size();
    h.depth4 = depth4;
    h.wide4.assign(32 * kids.size(), 0.0f);
    for (size_t k = 0; k < kids.size(); ++k) {
      float *o = &h.wide4[32 * k];
      const Kids &ch = kids[k];
      for (size_t q = 0; q < ch.size(); ++q) {
        const BuildNode &cn = B.nodes[ch[q]];
        for (int a = 0; a < 3; ++a) {
          o[4 * a + q] = cn.box.lo[a] - pad;
          o[4 * (3 + a) + q] = cn.box.hi[a] + pad;
        }
        int32_t ref;
        if (cn.left >= 0) {
          ref = n4_of[ch[q]];
        } else {
          if (cn.count >= 128) throw Error(ILLUM_ERR_CAPACITY, "leaf too large");
          ref = ~((cn.count << 24) | cn.first);
        }
        o[24 + q] = as_float(ref);
      }
      o[28] = as_float((int32_t)ch.size());
    }
  }
  for (int64_t k = 0; k < nf; ++k) {
    const int64_t t = f0 + B.perm[k];
    float *o = &h.tris[12 * (size_t)k];
    for (int v = 0; v < 3; ++v) {
      o[4 * v + 0] = g.tri[9 * t + 3 * v + 0];
      o[4 * v + 1] = g.tri[9 * t + 3 * v + 1];
      o[4 * v + 2] = g.tri[9 * t + 3 * v + 2];
      o[4 * v + 3] = v == 0 ? as_float((int32_t)t) : 0.0f;
    }
  }
  return h;
}

static inline int32_t as_int(float f) { int32_t i; std::memcpy(&i, &f, 4); return i; }

HostBVH build_bvh(const HostGeometry &g, int leaf_size) {
  if (leaf_size < 1 || leaf_size > 64) throw Error(ILLUM_ERR_ARG, "leaf_size must be in [1, 64]");
  const int64_t nf = g.nf;
  float scene_max = 0.0f;
  for (int64_t k = 0; k < 9 * nf; ++k) scene_max = std::max(scene_max, std::fabs(g.tri[k]));
  const float pad = scene_max * 4e-6f + FLT_MIN;
  std::vector<int64_t> off = g.body_off;
  if (off.size() < 2) off = {0, nf};
  const int nb = (int)off.size() - 1;
  if (nb == 1) {
    HostBVH h = build_one(g, 0, nf, leaf_size, pad);
    h.root_thr = {0};
    h.root_wide = {0};
    h.root_wide4 = {0};
    return h;
  }
  // Forest: the bodies' trees are built independently (in parallel, largest
  // first), then concatenated in body order, shifting node links and leaf
  // ranges.
  std::vector<HostBVH> trees(nb);
  {
    std::vector<int> order(nb);
    for (int b = 0; b < nb; ++b) order[b] = b;
    std::sort(order.begin(), order.end(),
              [&](int a, int c) { return off[a + 1] - off[a] > off[c + 1] - off[c]; });
    std::atomic<int> next(0);
    std::vector<std::string> errs(nb);
    auto worker = [&]() {
      for (int k; (k = next.fetch_add(1)) < nb;) {
        const int b = order[k];
        try {
          trees[b] = build_one(g, off[b], off[b + 1], leaf_size, pad);
        } catch (const std::exception &e) {
          errs[b] = e.what();
        }
      }
    };
    const int nt = std::max(1, std::min<int>(nb, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
    for (int b = 0; b < nb; ++b)
      if (!errs[b].empty()) throw Error(ILLUM_ERR_ARG, "BVH of body " + std::to_string(b) + ": " + errs[b]);
  }
  HostBVH h;
  h.leaf_size = leaf_size;
  h.tris.resize(12 * (size_t)nf);
  for (int b = 0; b < nb; ++b) {
    HostBVH t = std::move(trees[b]);
    const int32_t nbase = (int32_t)h.nnodes, ibase = (int32_t)h.ninner;
    const int32_t qbase = (int32_t)h.ninner4;
    const int32_t tbase = (int32_t)off[b];
    if (off[b + 1] >= (1 << 24)) throw Error(ILLUM_ERR_CAPACITY, "too many facets for leaf ranges");
    h.root_thr.push_back(nbase);
    h.root_wide.push_back(ibase);
    h.root_wide4.push_back(qbase);
    h.depth = std::max(h.depth, t.depth);
    h.depth4 = std::max(h.depth4, t.depth4);
    for (int64_t k = 0; k < t.ninner4; ++k) {
      float *o = &t.wide4[32 * (size_t)k];
      const int32_t cnt = as_int(o[28]);
      for (int q = 0; q < cnt; ++q) {
        const int32_t ref = as_int(o[24 + q]);
        o[24 + q] = as_float(ref >= 0 ? ref + qbase : ~(~ref + tbase));
      }
    }
    h.wide4.insert(h.wide4.end(), t.wide4.begin(), t.wide4.end());
    h.ninner4 += t.ninner4;
    for (int64_t k = 0; k < t.nnodes; ++k) {
      float *o = &t.nodes[8 * (size_t)k];
      const int32_t miss = as_int(o[3]), info = as_int(o[7]);
      o[3] = as_float(miss < 0 ? -1 : miss + nbase);
      if (info >= 0) o[7] = as_float(info + tbase);  // count << 24 | first
    }
    for (int64_t k = 0; k < 2 * t.ninner; ++k) {
      float *o = &t.wide[8 * (size_t)k];
      const int32_t ref = as_int(o[3]);
      o[3] = as_float(ref >= 0 ? ref + ibase : ~(~ref + tbase));
    }
    h.nodes.insert(h.nodes.end(), t.nodes.begin(), t.nodes.end());
    h.wide.insert(h.wide.end(), t.wide.begin(), t.wide.end());
    std::copy(t.tris.begin(), t.tris.end(), h.tris.begin() + 12 * (size_t)off[b]);
    h.nnodes += t.nnodes;
    h.ninner += t.ninner;
  }
  return h;
}

}  // namespace illum
