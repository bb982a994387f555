// CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++ restatement of the reference hot path, used by tests/, by
// __graft_entry__.smoke() and by bench.py's CPU-baseline leg as the checker.
// The product library never links, loads or calls anything in this directory.
// It shares no source with the product: geometry, BVH, traversal and all
// numerics are written again here from the reference and the published
// methods it names.
//
// Functions and the reference code they restate (paths relative to the illum
// source tree):
//   oracle_geometry          obj_util.cpp:43-93 (face normal, /100), illum.cpp:101-105
//                            (area), illum.cpp:244-262 (Frenet frame)
//   oracle_closest_hit       the FP32 watertight predicate (Woop, Benthin & Wald,
//                            JCGT 2013) used for every visibility ray; brute force
//                            over all triangles or an independent median-split BVH
//   oracle_visibility        compute_F's visibility loop, illum.cpp:42-95
//   oracle_form_factors      compute_F values + assembly, illum.cpp:100-234
//   oracle_horizons          trace_horizon / make_horizons, illum.cpp:274-363
//   oracle_fib_spiral        fib_spiral, illum.cpp:468-483
//   oracle_direct            get_direct_radiosity, illum.cpp:382-466 (deterministic
//                            disk tangent instead of randn, illum.cpp:406)
//   oracle_gauss_seidel      illum_cli.cpp:425-440 + arma_util.cpp:3-19
//   oracle_conductionQ       conductionQ.f:35-98 with the Thomas algorithm of
//                            tridag.for:17-27 (textbook form, Golub & Van Loan 4.3)
//   oracle_thermal_run       thermal_model::step, thermal.hpp:65-84
//   oracle_render            the renderer's per-pixel ray-cast, illum_renderer_cli.cpp:174-261
//   oracle_walk_work         measurement aid, no reference counterpart: the box and
//                            triangle tests of compute_F's rays (illum.cpp:56-73) along
//                            the product's own exported tree, walked one ray at a time
//
// Compiled with -ffp-contract=off: every expression is evaluated as written.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

extern "C" {

// --------------------------------------------------------------------------
// Geometry.
int oracle_geometry(const double *V, int64_t nv, const int32_t *F, int64_t nf,
                    const double *VN, double divisor, float *tri9, float *cen, float *nrm,
                    float *rad, double *area, double *frame) {
  if (!V || !F || nf <= 0) return 1;
  for (int64_t f = 0; f < nf; ++f) {
    // w: the FP32 vertices as loaded; v = w / divisor.  The normal is taken
    // from w (obj_util.cpp:70), the triangle from v (obj_util.cpp:92).
    float w[3][3], v[3][3];
    for (int k = 0; k < 3; ++k) {
      const int32_t id = F[3 * f + k];
      if (id < 0 || id >= nv) return 1;
      for (int c = 0; c < 3; ++c) {
        w[k][c] = (float)V[3 * id + c];
        v[k][c] = divisor == 1.0 ? w[k][c] : w[k][c] / (float)divisor;
      }
    }
    if (tri9)
      for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) tri9[9 * f + 3 * k + c] = v[k][c];
    // n = normalize((w1 - w0) x (w2 - w0))
    const float ax = w[1][0] - w[0][0], ay = w[1][1] - w[0][1], az = w[1][2] - w[0][2];
    const float bx = w[2][0] - w[0][0], by = w[2][1] - w[0][1], bz = w[2][2] - w[0][2];
    const float cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    const float len = std::sqrt((cx * cx + cy * cy) + cz * cz);
    float n0 = cx / len, n1 = cy / len, n2 = cz / len;
    if (VN) {
      float s[3];
      for (int c = 0; c < 3; ++c)
        s[c] = ((float)VN[3 * F[3 * f] + c] + (float)VN[3 * F[3 * f + 1] + c]) +
               (float)VN[3 * F[3 * f + 2] + c];
      if ((n0 * s[0] + n1 * s[1]) + n2 * s[2] < 0) { n0 = -n0; n1 = -n1; n2 = -n2; }
    }
    float p[3];
    for (int c = 0; c < 3; ++c) p[c] = ((v[0][c] + v[1][c]) + v[2][c]) / 3.0f;
    float r = 0.0f;
    for (int k = 0; k < 3; ++k) {
      const float dx = v[k][0] - p[0], dy = v[k][1] - p[1], dz = v[k][2] - p[2];
      const float d = std::sqrt((dx * dx + dy * dy) + dz * dz);
      if (d > r) r = d;
    }
    if (cen) { cen[3 * f] = p[0]; cen[3 * f + 1] = p[1]; cen[3 * f + 2] = p[2]; }
    if (nrm) { nrm[3 * f] = n0; nrm[3 * f + 1] = n1; nrm[3 * f + 2] = n2; }
    if (rad) rad[f] = r;
    if (area) {
      double e1[3], e2[3];
      for (int c = 0; c < 3; ++c) {
        e1[c] = (double)v[1][c] - (double)v[0][c];
        e2[c] = (double)v[2][c] - (double)v[0][c];
      }
      const double qx = e1[1] * e2[2] - e1[2] * e2[1], qy = e1[2] * e2[0] - e1[0] * e2[2],
                   qz = e1[0] * e2[1] - e1[1] * e2[0];
      area[f] = std::sqrt((qx * qx + qy * qy) + qz * qz) / 2.0;
    }
    if (frame) {
      const double nx = n0, ny = n1, nz = n2;
      const double nn = std::sqrt((nx * nx + ny * ny) + nz * nz);
      const double ux = nx / nn, uy = ny / nn, uz = nz / nn;
      const double tx = (double)(v[1][0] - v[0][0]), ty = (double)(v[1][1] - v[0][1]),
                   tz = (double)(v[1][2] - v[0][2]);
      const double tn = std::sqrt((tx * tx + ty * ty) + tz * tz);
      const double t0 = tx / tn, t1 = ty / tn, t2 = tz / tn;
      double *o = frame + 9 * f;
      o[0] = t1 * uz - t2 * uy; o[1] = t2 * ux - t0 * uz; o[2] = t0 * uy - t1 * ux;
      o[3] = t0; o[4] = t1; o[5] = t2;
      o[6] = ux; o[7] = uy; o[8] = uz;
    }
  }
  return 0;
}

}  // extern "C"

namespace {

// --------------------------------------------------------------------------
// FP32 watertight ray/triangle predicate.
struct Ray {
  float o[3];
  float inv[3];
  float no[3];  // -o * inv: slab function s(x) = fma(x, inv, no)
  float S[3];
  int k[3];  // kx, ky, kz
};

inline float axis_inv(float d) {
  if (d == 0.0f) d = std::signbit(d) ? -1e-30f : 1e-30f;
  return 1.0f / d;
}

Ray setup(const float *o, const float *d) {
  Ray r;
  for (int c = 0; c < 3; ++c) r.o[c] = o[c];
  int kz = 0;
  float m = std::fabs(d[0]);
  if (std::fabs(d[1]) > m) { kz = 1; m = std::fabs(d[1]); }
  if (std::fabs(d[2]) > m) kz = 2;
  int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
  if (d[kz] < 0.0f) std::swap(kx, ky);
  r.k[0] = kx; r.k[1] = ky; r.k[2] = kz;
  r.S[0] = d[kx] / d[kz];
  r.S[1] = d[ky] / d[kz];
  r.S[2] = 1.0f / d[kz];
  for (int c = 0; c < 3; ++c) r.inv[c] = axis_inv(d[c]);
  for (int c = 0; c < 3; ++c) r.no[c] = -(r.o[c] * r.inv[c]);
  return r;
}

// Returns true on a hit with t > 0; *t = max(t, entry into the triangle's box).
bool hit_tri(const Ray &r, const float *v0, const float *v1, const float *v2, float *tout) {
  float A[3], B[3], C[3];
  for (int c = 0; c < 3; ++c) {
    A[c] = v0[c] - r.o[c];
    B[c] = v1[c] - r.o[c];
    C[c] = v2[c] - r.o[c];
  }
  const int kx = r.k[0], ky = r.k[1], kz = r.k[2];
  const float Ax = A[kx] - r.S[0] * A[kz], Ay = A[ky] - r.S[1] * A[kz];
  const float Bx = B[kx] - r.S[0] * B[kz], By = B[ky] - r.S[1] * B[kz];
  const float Cx = C[kx] - r.S[0] * C[kz], Cy = C[ky] - r.S[1] * C[kz];
  float U = Cx * By - Cy * Bx;
  float Vv = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || Vv == 0.0f || W == 0.0f) {
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    Vv = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || Vv < 0.0f || W < 0.0f) && (U > 0.0f || Vv > 0.0f || W > 0.0f)) return false;
  const float det = (U + Vv) + W;
  if (det == 0.0f) return false;
  const float Az = r.S[2] * A[kz], Bz = r.S[2] * B[kz], Cz = r.S[2] * C[kz];
  const float T = (U * Az + Vv * Bz) + W * Cz;
  const float t = T / det;
  if (!(t > 0.0f)) return false;
  float e = -FLT_MAX;
  for (int c = 0; c < 3; ++c) {
    const float lo = std::fmin(std::fma(v0[c], r.inv[c], r.no[c]),
                               std::fmin(std::fma(v1[c], r.inv[c], r.no[c]),
                                         std::fma(v2[c], r.inv[c], r.no[c])));
    e = std::fmax(e, lo);
  }
  *tout = std::fmax(t, e);
  return true;
}

// Independent binary BVH (object-median split on the widest centroid axis,
// explicit stack, near child first).  Boxes padded by 1e-5 of the scene's
// largest coordinate.
struct BVH {
  struct Node { float lo[3], hi[3]; int left, right, first, count; };
  std::vector<Node> nodes;
  std::vector<int> idx;
  const float *tri = nullptr;
  int64_t nf = 0;
  float pad = 0;

  void build(const float *tri9, int64_t n) {
    tri = tri9;
    nf = n;
    idx.resize(n);
    std::vector<float> cen(3 * n);
    float smax = 0;
    for (int64_t t = 0; t < n; ++t) {
      idx[t] = (int)t;
      for (int c = 0; c < 3; ++c) {
        float lo = tri9[9 * t + c], hi = lo;
        for (int k = 1; k < 3; ++k) {
          lo = std::min(lo, tri9[9 * t + 3 * k + c]);
          hi = std::max(hi, tri9[9 * t + 3 * k + c]);
        }
        cen[3 * t + c] = 0.5f * (lo + hi);
        smax = std::max(smax, std::max(std::fabs(lo), std::fabs(hi)));
      }
    }
    pad = smax * 1e-5f + FLT_MIN;
    nodes.clear();
    nodes.reserve(2 * n);
    rec(0, (int)n, cen);
  }

  int rec(int first, int count, const std::vector<float> &cen) {
    const int id = (int)nodes.size();
    nodes.push_back(Node());
    Node nd;
    for (int c = 0; c < 3; ++c) { nd.lo[c] = FLT_MAX; nd.hi[c] = -FLT_MAX; }
    float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int q = first; q < first + count; ++q) {
      const int t = idx[q];
      for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) {
          nd.lo[c] = std::min(nd.lo[c], tri[9 * t + 3 * k + c]);
          nd.hi[c] = std::max(nd.hi[c], tri[9 * t + 3 * k + c]);
        }
      for (int c = 0; c < 3; ++c) {
        clo[c] = std::min(clo[c], cen[3 * t + c]);
        chi[c] = std::max(chi[c], cen[3 * t + c]);
      }
    }
    for (int c = 0; c < 3; ++c) { nd.lo[c] -= pad; nd.hi[c] += pad; }
    nd.left = nd.right = -1;
    nd.first = first;
    nd.count = count;
    if (count > 4) {
      int ax = 0;
      for (int c = 1; c < 3; ++c)
        if (chi[c] - clo[c] > chi[ax] - clo[ax]) ax = c;
      const int mid = first + count / 2;
      std::nth_element(idx.begin() + first, idx.begin() + mid, idx.begin() + first + count,
                       [&](int a, int b) {
                         return cen[3 * a + ax] < cen[3 * b + ax] ||
                                (cen[3 * a + ax] == cen[3 * b + ax] && a < b);
                       });
      nodes[id] = nd;
      const int l = rec(first, mid - first, cen);
      const int r = rec(mid, first + count - mid, cen);
      nodes[id].left = l;
      nodes[id].right = r;
      return id;
    }
    nodes[id] = nd;
    return id;
  }

  static bool box(const Ray &r, const Node &n, float tbest, float *tn_out) {
    float tn = -FLT_MAX, tf = FLT_MAX;
    for (int c = 0; c < 3; ++c) {
      const float t0 = std::fma(n.lo[c], r.inv[c], r.no[c]);
      const float t1 = std::fma(n.hi[c], r.inv[c], r.no[c]);
      tn = std::fmax(tn, std::fmin(t0, t1));
      tf = std::fmin(tf, std::fmax(t0, t1));
    }
    *tn_out = tn;
    return tn <= tf && tf > 0.0f && tn <= tbest;
  }

  int closest(const Ray &r, float *tb_out, bool any) const {
    float tbest = HUGE_VALF;
    int best = -1;
    int stack[128];
    int sp = 0;
    float tn;
    if (!box(r, nodes[0], tbest, &tn)) { *tb_out = tbest; return -1; }
    stack[sp++] = 0;
    while (sp) {
      const Node &n = nodes[stack[--sp]];
      float tchk;
      if (!box(r, n, tbest, &tchk)) continue;
      if (n.left < 0) {
        for (int q = n.first; q < n.first + n.count; ++q) {
          const int t = idx[q];
          float th;
          if (hit_tri(r, tri + 9 * t, tri + 9 * t + 3, tri + 9 * t + 6, &th)) {
            if (any) { *tb_out = th; return t; }
            if (th < tbest || (th == tbest && t < best)) { tbest = th; best = t; }
          }
        }
        continue;
      }
      float tl, tr;
      const bool hl = box(r, nodes[n.left], tbest, &tl);
      const bool hr = box(r, nodes[n.right], tbest, &tr);
      if (hl && hr) {
        if (tl <= tr) { stack[sp++] = n.right; stack[sp++] = n.left; }
        else { stack[sp++] = n.left; stack[sp++] = n.right; }
      } else if (hl) {
        stack[sp++] = n.left;
      } else if (hr) {
        stack[sp++] = n.right;
      }
    }
    *tb_out = tbest;
    return best;
  }
};

int brute_closest(const float *tri, int64_t nf, const Ray &r, float *tb_out) {
  float tbest = HUGE_VALF;
  int best = -1;
  for (int64_t t = 0; t < nf; ++t) {
    float th;
    if (hit_tri(r, tri + 9 * t, tri + 9 * t + 3, tri + 9 * t + 6, &th) &&
        (th < tbest || (th == tbest && (int)t < best))) {
      tbest = th;
      best = (int)t;
    }
  }
  *tb_out = tbest;
  return best;
}

template <class Fn>
void parallel_for(int64_t n, int nthreads, Fn fn) {
  if (nthreads <= 1 || n < 2) {
    for (int64_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<int64_t> next(0);
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&] {
      for (;;) {
        const int64_t i = next.fetch_add(1);
        if (i >= n) break;
        fn(i);
      }
    });
  for (auto &x : th) x.join();
}

// Form factor F(i, j): receiver i, source j (illum.cpp:100-119) in FP64 from
// the FP32 centroids and normals.
double ff_value(const float *pi, const float *ni, const float *pj, const float *nj, double Aj) {
  const double px = (double)pj[0] - (double)pi[0];
  const double py = (double)pj[1] - (double)pi[1];
  const double pz = (double)pj[2] - (double)pi[2];
  const double r = std::sqrt((px * px + py * py) + pz * pz);
  if (r < DBL_EPSILON) return 0.0;
  const double ux = px / r, uy = py / r, uz = pz / r;
  const double mij = std::fmax(0.0, ((double)ni[0] * ux + (double)ni[1] * uy) + (double)ni[2] * uz);
  const double mji = std::fmax(0.0, -(((double)nj[0] * ux + (double)nj[1] * uy) + (double)nj[2] * uz));
  return ((mij * mji) * Aj) / ((M_PI * r) * r);
}

}  // namespace

// ---- Work of the product's per-ray walk ----------------------------------
// The form-factor rays of sources [j0, j1) (those of oracle_visibility) walked
// one at a time over the product library's own tree, exported as its walks
// read it (illum_bvh_export), counting box and triangle tests: the near-first
// walk of the two-wide tree (visit the child entered first, push the other,
// prune boxes by the best distance so far), first bounded at t <= tinit and,
// if it found nothing, unbounded; a walk that would push more than cap
// entries ends there and the stackless walk of the threaded tree answers the
// unbounded query instead.  These counts price a ray's own walk (12 FLOP per
// box test, 40 per triangle test) for the visibility kernel's roofline.
// out = {rays, box tests, triangle tests}; bits (nullable) receives the
// closest hits as in oracle_visibility.
namespace {

inline int as_int(float f) {
  int i;
  std::memcpy(&i, &f, 4);
  return i;
}

// Slab test of the box lo[0..2], hi[0..2]: entry distance tn; accepted when
// tn <= tf, tf > 0 and tn <= tbest (max(tn, m) <= min(tf, tbest), m the
// smallest positive float).
bool slab_enter(const Ray &r, const float *lo, const float *hi, float tbest, float *tn_out) {
  float tn = -INFINITY, tf = INFINITY;
  for (int c = 0; c < 3; ++c) {
    const float t0 = std::fma(lo[c], r.inv[c], r.no[c]), t1 = std::fma(hi[c], r.inv[c], r.no[c]);
    tn = std::fmax(tn, std::fmin(t0, t1));
    tf = std::fmin(tf, std::fmax(t0, t1));
  }
  *tn_out = tn;
  return std::fmax(tn, std::numeric_limits<float>::denorm_min()) <= std::fmin(tf, tbest);
}

struct ExportedTree {
  const float *nodes, *wide, *tris;
  int root_thr, root_wide, cap;
};

void leaf_tests(const ExportedTree &w, const Ray &r, int info, float &tbest, int &best,
                int64_t &ntri) {
  const int first = info & 0xFFFFFF, cnt = info >> 24;
  ntri += cnt;
  for (int k = 0; k < cnt; ++k) {
    const float *t = w.tris + 12 * (int64_t)(first + k);
    const int idx = as_int(t[3]);
    float d;
    if (hit_tri(r, t, t + 4, t + 8, &d) && (d < tbest || (d == tbest && idx < best))) {
      tbest = d;
      best = idx;
    }
  }
}

// Near-first walk of the two-wide tree: the facet hit, -1 for none, -2 when
// the walk needed more than cap stack entries.
int walk_two_wide(const ExportedTree &w, const Ray &r, float tinit, int64_t &nbox,
                  int64_t &ntri) {
  float tbest = tinit;
  int best = -1, sp = 0, cur = w.root_wide;
  int stack[64];
  for (;;) {
    if (cur >= 0) {
      nbox += 2;
      const float *nd = w.wide + 16 * (int64_t)cur;
      float tl, tr;
      const bool hl = slab_enter(r, nd, nd + 4, tbest, &tl);
      const bool hr = slab_enter(r, nd + 8, nd + 12, tbest, &tr);
      const int cl = as_int(nd[3]), cr = as_int(nd[11]);
      if (hl && hr) {
        if (sp == w.cap || sp == 64) return -2;
        const bool lnear = tl <= tr;
        stack[sp++] = lnear ? cr : cl;
        cur = lnear ? cl : cr;
        continue;
      }
      if (hl) { cur = cl; continue; }
      if (hr) { cur = cr; continue; }
    } else {
      leaf_tests(w, r, ~cur, tbest, best, ntri);
    }
    if (sp == 0) break;
    cur = stack[--sp];
  }
  return best;
}

// Stackless walk of the threaded tree (miss links), unbounded.
int walk_threaded(const ExportedTree &w, const Ray &r, int64_t &nbox, int64_t &ntri) {
  float tbest = INFINITY;
  int best = -1, node = w.root_thr;
  while (node >= 0) {
    const float *nd = w.nodes + 8 * (int64_t)node;
    const int miss = as_int(nd[3]), info = as_int(nd[7]);
    ++nbox;
    float tn;
    if (!slab_enter(r, nd, nd + 4, tbest, &tn)) { node = miss; continue; }
    if (info < 0) { node = node + 1; continue; }
    leaf_tests(w, r, info, tbest, best, ntri);
    node = miss;
  }
  return best;
}

}  // namespace

extern "C" {

int oracle_walk_work(const float *cen, const float *nrm, const float *rad, int64_t nf,
                     double offset, int64_t j0, int64_t j1, const float *nodes, int root_thr,
                     const float *wide, int root_wide, const float *tris, int cap, float tinit,
                     int nthreads, int64_t *out, uint32_t *bits) {
  const ExportedTree w{nodes, wide, tris, root_thr, root_wide, cap};
  const int64_t words = (nf + 31) / 32;
  std::atomic<int64_t> nrays(0), nbox(0), ntri(0);
  const float off = (float)offset;
  parallel_for(j1 - j0, nthreads, [&](int64_t q) {
    const int64_t j = j0 + q;
    uint32_t *row = bits ? bits + q * words : nullptr;
    if (row) std::memset(row, 0, 4 * words);
    const float *pj = cen + 3 * j, *nj = nrm + 3 * j;
    const float o[3] = {pj[0] + off * nj[0], pj[1] + off * nj[1], pj[2] + off * nj[2]};
    int64_t lr = 0, lb = 0, lt = 0;
    for (int64_t i = 0; i < nf; ++i) {
      if (i == j) continue;
      const float *pi = cen + 3 * i;
      const float dx = pi[0] - pj[0], dy = pi[1] - pj[1], dz = pi[2] - pj[2];
      if ((dx * nj[0] + dy * nj[1]) + dz * nj[2] <= -rad[i]) continue;
      const float d[3] = {pi[0] - o[0], pi[1] - o[1], pi[2] - o[2]};
      const Ray r = setup(o, d);
      ++lr;
      int h = walk_two_wide(w, r, tinit, lb, lt);
      if (h == -1) h = walk_two_wide(w, r, INFINITY, lb, lt);
      if (h == -2) h = walk_threaded(w, r, lb, lt);
      if (row && h >= 0 && h != j) row[h >> 5] |= 1u << (h & 31);
    }
    nrays += lr;
    nbox += lb;
    ntri += lt;
  });
  out[0] = nrays.load();
  out[1] = nbox.load();
  out[2] = ntri.load();
  return 0;
}

}  // extern "C"

extern "C" {

// Closest hit for nray rays (origins/directions nray x 3) over nf triangles
// (nf x 9 floats).  use_bvh = 0: brute force; 1: the oracle's own BVH.
int oracle_closest_hit(const float *tri9, int64_t nf, const float *orig, const float *dir,
                       int64_t nray, int use_bvh, int nthreads, int32_t *hit, float *t) {
  BVH bvh;
  if (use_bvh) bvh.build(tri9, nf);
  parallel_for(nray, nthreads, [&](int64_t k) {
    const Ray r = setup(orig + 3 * k, dir + 3 * k);
    float tb;
    hit[k] = use_bvh ? bvh.closest(r, &tb, false) : brute_closest(tri9, nf, r, &tb);
    t[k] = tb;
  });
  return 0;
}

// Visibility sets of compute_F (illum.cpp:42-95) for sources [j0, j1):
// bit i of row (j - j0) set iff facet i is the closest hit of a ray from
// q_j = p_j + offset n_j towards the centroid of an un-culled facet.
// words = ceil(nf / 32) words per row.  Returns the number of rays cast.
int64_t oracle_visibility(const float *tri9, const float *cen, const float *nrm,
                          const float *rad, int64_t nf, double offset, int64_t j0, int64_t j1,
                          int use_bvh, int nthreads, uint32_t *bits) {
  BVH bvh;
  if (use_bvh) bvh.build(tri9, nf);
  const int64_t words = (nf + 31) / 32;
  std::atomic<int64_t> nrays(0);
  const float off = (float)offset;
  parallel_for(j1 - j0, nthreads, [&](int64_t q) {
    const int64_t j = j0 + q;
    uint32_t *row = bits + q * words;
    std::memset(row, 0, 4 * words);
    const float *pj = cen + 3 * j, *nj = nrm + 3 * j;
    const float o[3] = {pj[0] + off * nj[0], pj[1] + off * nj[1], pj[2] + off * nj[2]};
    int64_t local = 0;
    for (int64_t i = 0; i < nf; ++i) {
      if (i == j) continue;
      const float *pi = cen + 3 * i;
      const float dx = pi[0] - pj[0], dy = pi[1] - pj[1], dz = pi[2] - pj[2];
      if ((dx * nj[0] + dy * nj[1]) + dz * nj[2] <= -rad[i]) continue;
      const float d[3] = {pi[0] - o[0], pi[1] - o[1], pi[2] - o[2]};
      const Ray r = setup(o, d);
      float tb;
      const int h = use_bvh ? bvh.closest(r, &tb, false) : brute_closest(tri9, nf, r, &tb);
      ++local;
      if (h < 0 || h == j) continue;
      row[h >> 5] |= 1u << (h & 31);
    }
    nrays += local;
  });
  return nrays.load();
}

// Form factors for sources [j0, j1) from visibility rows (as produced by
// oracle_visibility): CSC column pointers relative to j0 (colptr[0] = 0),
// sorted row indices and FP64 values scaled by albedo (albedo_vec if non-null),
// exact zeros dropped.  Returns nnz, -1 if cap is too small, -2 on a NaN.
int64_t oracle_form_factors(const float *cen, const float *nrm, const double *area, int64_t nf,
                            const uint32_t *bits, int64_t j0, int64_t j1, double albedo,
                            const double *albedo_vec, int64_t cap, int64_t *colptr,
                            int32_t *rowidx, double *val) {
  const int64_t words = (nf + 31) / 32;
  int64_t nnz = 0;
  colptr[0] = 0;
  for (int64_t j = j0; j < j1; ++j) {
    const uint32_t *row = bits + (j - j0) * words;
    for (int64_t w = 0; w < words; ++w) {
      uint32_t x = row[w];
      while (x) {
        const int b = __builtin_ctz(x);
        x &= x - 1;
        const int64_t i = 32 * w + b;
        const double rho = albedo_vec ? albedo_vec[i] : albedo;
        const double v = ff_value(cen + 3 * i, nrm + 3 * i, cen + 3 * j, nrm + 3 * j, area[j]) * rho;
        if (std::isnan(v)) return -2;
        if (v == 0.0) continue;
        if (nnz >= cap) return -1;
        rowidx[nnz] = (int32_t)i;
        val[nnz] = v;
        ++nnz;
      }
    }
    colptr[j - j0 + 1] = nnz;
  }
  return nnz;
}

// Horizon map (illum.cpp:274-363) for facets [j0, j1): out is
// nphi x (j1 - j0) column-major.  Occluder triangles occ9 (nocc facets) may
// differ from the facets' own mesh.  Returns the number of rays cast.
int64_t oracle_horizons(const float *cen, const float *nrm, const double *frame, int64_t j0,
                        int64_t j1, const float *occ9, int64_t nocc, int nphi, double theta_eps,
                        double offset, int use_bvh, int nthreads, double *out) {
  BVH bvh;
  if (use_bvh) bvh.build(occ9, nocc);
  std::vector<double> phis(nphi);
  const double dphi = 2 * M_PI / (nphi - 1);
  for (int k = 0; k < nphi; ++k) phis[k] = k == nphi - 1 ? 2 * M_PI : k * dphi;
  std::atomic<int64_t> nrays(0);
  const float off = (float)offset;
  parallel_for(j1 - j0, nthreads, [&](int64_t q) {
    const int64_t j = j0 + q;
    const float *p = cen + 3 * j, *n = nrm + 3 * j;
    const float o[3] = {p[0] + off * n[0], p[1] + off * n[1], p[2] + off * n[2]};
    const double *F = frame + 9 * j;
    int64_t local = 0;
    for (int k = 0; k < nphi; ++k) {
      const double ph = phis[k];
      double lo = 0, hi = M_PI, mid = hi / 2, prev;
      do {
        prev = mid;
        const double vx = std::cos(ph) * std::sin(mid), vy = std::sin(ph) * std::sin(mid),
                     vz = std::cos(mid);
        const float d[3] = {(float)((F[0] * vx + F[3] * vy) + F[6] * vz),
                            (float)((F[1] * vx + F[4] * vy) + F[7] * vz),
                            (float)((F[2] * vx + F[5] * vy) + F[8] * vz)};
        const Ray r = setup(o, d);
        float tb;
        const bool hit = use_bvh ? bvh.closest(r, &tb, true) >= 0 : brute_closest(occ9, nocc, r, &tb) >= 0;
        ++local;
        if (hit) hi = mid;
        else lo = mid;
        mid = (lo + hi) / 2;
      } while (std::fabs(mid - prev) > theta_eps);
      out[q * nphi + k] = mid;
    }
    nrays += local;
  });
  return nrays.load();
}

int oracle_fib_spiral(double *xy, int n) {
  const double TWO_PI = 2 * M_PI, gr = 1.6180339887498948482;
  double a = 0., da = (TWO_PI * (gr - 1)) / gr;
  double r = 0., dr = 1. / (n + 1);
  for (int i = 0; i < n; ++i) {
    xy[i] = r * std::cos(a);
    xy[n + i] = r * std::sin(a);
    a = std::fmod(a + da, TWO_PI);
    r += dr;
  }
  return 0;
}

// Direct insolation (illum.cpp:382-466) for all facets and S sun positions
// (km, 3 x S column-major).  horizons: nphi x nf; disk: ndisk x 2 (col-major).
// E: nf x S column-major.  The random disk tangent of illum.cpp:406 is
// replaced by normalise((I - N N^T) e_k), e_k the axis least aligned with N.
int oracle_direct(const float *cen, const double *frame, int64_t nf, const double *hz, int nphi,
                  const double *sun, int S, const double *disk, int ndisk, int nthreads,
                  double *E) {
  const double TWO_PI = 2 * M_PI, AU = 149.5978707e6, RS = 695508;
  const double dphi = TWO_PI / (nphi - 1);
  parallel_for(nf * S, nthreads, [&](int64_t idx) {
    const int64_t f = idx % nf;
    const int s = (int)(idx / nf);
    const double p[3] = {cen[3 * f], cen[3 * f + 1], cen[3 * f + 2]};
    const double *sp = sun + 3 * s;
    double N[3] = {sp[0] - p[0], sp[1] - p[1], sp[2] - p[2]};
    const double nN = std::sqrt((N[0] * N[0] + N[1] * N[1]) + N[2] * N[2]);
    for (int c = 0; c < 3; ++c) N[c] = N[c] / nN;
    int e = 0;
    double mn = std::fabs(N[0]);
    if (std::fabs(N[1]) < mn) { e = 1; mn = std::fabs(N[1]); }
    if (std::fabs(N[2]) < mn) e = 2;
    double T[3];
    for (int c = 0; c < 3; ++c) T[c] = (c == e ? 1.0 : 0.0) - N[c] * N[e];
    const double nT = std::sqrt((T[0] * T[0] + T[1] * T[1]) + T[2] * T[2]);
    for (int c = 0; c < 3; ++c) T[c] = T[c] / nT;
    const double B[3] = {T[1] * N[2] - T[2] * N[1], T[2] * N[0] - T[0] * N[2],
                         T[0] * N[1] - T[1] * N[0]};
    const double *F = frame + 9 * f;
    const double *h = hz + f * nphi;
    int count = 0;
    for (int k = 0; k < ndisk - 1; ++k) {
      const double xk = disk[k], yk = disk[ndisk + k];
      double w[3];
      for (int c = 0; c < 3; ++c) w[c] = (sp[c] + RS * (xk * T[c] + yk * B[c])) - p[c];
      const double nw = std::sqrt((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
      for (int c = 0; c < 3; ++c) w[c] = w[c] / nw;
      const double d0 = (F[0] * w[0] + F[1] * w[1]) + F[2] * w[2];
      const double d1 = (F[3] * w[0] + F[4] * w[1]) + F[5] * w[2];
      const double d2 = (F[6] * w[0] + F[7] * w[1]) + F[8] * w[2];
      const double theta = std::acos(d2);
      double phi = std::atan2(d1, d0);
      if (phi < 0) phi = phi + TWO_PI;
      int pk = (int)std::floor(phi / dphi);
      pk = pk < 0 ? 0 : (pk > nphi - 1 ? nphi - 1 : pk);
      const double t = (phi - pk * dphi) / dphi;
      if (theta < h[pk] + t * dphi) ++count;
    }
    double v = (double)count / (nphi - 1);
    v = v * std::fmax(0.0, (N[0] * F[6] + N[1] * F[7]) + N[2] * F[8]);
    const double ns = std::sqrt((sp[0] * sp[0] + sp[1] * sp[1]) + sp[2] * sp[2]);
    const double q = AU / ns;
    v = v * (1372 * (q * q));
    E[(int64_t)s * nf + f] = v;
  });
  return 0;
}

// Gauss-Seidel radiosity (illum_cli.cpp:425-440) on F in CSC (column j =
// source).  L = I - tril(F), U = -triu(F); B0 = L^-1 E, then up to max_sweeps
// more sweeps B <- L^-1 (E - U B), stopping early once the relative residual
// ||L B + U B - E|| / ||E|| <= rel_tol (rel_tol <= 0: run all sweeps, as the
// reference's fixed gs_steps).  abs_resid (nullable, max_sweeps entries)
// receives the reference's absolute residual after each sweep.  Returns the
// number of sweeps after B0.
int oracle_gauss_seidel(int64_t n, const int64_t *colptr, const int32_t *rowidx,
                        const double *val, const double *E, int max_sweeps, double rel_tol,
                        double *B, double *abs_resid) {
  // CSR of F (row i, ascending j) -- the traversal order of arma's column walk
  // of L^T = CSR of L.
  std::vector<int64_t> rp(n + 1, 0);
  const int64_t nnz = colptr[n];
  for (int64_t k = 0; k < nnz; ++k) rp[rowidx[k] + 1]++;
  for (int64_t i = 0; i < n; ++i) rp[i + 1] += rp[i];
  std::vector<int32_t> ci(nnz);
  std::vector<double> cv(nnz);
  std::vector<int64_t> cur(rp.begin(), rp.end() - 1);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k) {
      const int64_t i = rowidx[k];
      ci[cur[i]] = (int32_t)j;
      cv[cur[i]] = val[k];
      cur[i]++;
    }
  std::vector<double> rhs(E, E + n), UB(n);
  double enorm = 0;
  for (int64_t i = 0; i < n; ++i) enorm += E[i] * E[i];
  enorm = std::sqrt(enorm);
  // forward_solve(L_t, b): x_i = (b_i - sum_{j<i} L_ij x_j) / L_ii, L_ij = -F_ij, L_ii = 1.
  auto forward = [&](const std::vector<double> &b) {
    for (int64_t i = 0; i < n; ++i) {
      double x = b[i];
      double diag = 1.0;
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
        const int64_t j = ci[k];
        if (j < i) x -= (-cv[k]) * B[j];
        else if (j == i) diag = 1.0 - cv[k];
      }
      B[i] = x / diag;
    }
  };
  auto upper = [&]() {  // UB = U B = -triu(F) B
    for (int64_t i = 0; i < n; ++i) {
      double y = 0;
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] > i) y += (-cv[k]) * B[ci[k]];
      UB[i] = y;
    }
  };
  auto residual = [&]() {  // ||L B + U B - E||
    double s = 0;
    for (int64_t i = 0; i < n; ++i) {
      double lb = 0;
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] < i) lb += (-cv[k]) * B[ci[k]];
        else if (ci[k] == i) lb += (1.0 - cv[k]) * B[i];
      if (std::find(ci.begin() + rp[i], ci.begin() + rp[i + 1], (int32_t)i) == ci.begin() + rp[i + 1])
        lb += B[i];
      double ub = 0;
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] > i) ub += (-cv[k]) * B[ci[k]];
      const double r = (lb + ub) - E[i];
      s += r * r;
    }
    return std::sqrt(s);
  };
  forward(rhs);
  int sweeps = 0;
  for (int it = 0; it < max_sweeps; ++it) {
    if (rel_tol > 0) {
      const double r = residual();
      if (r <= rel_tol * enorm) break;
    }
    upper();
    for (int64_t i = 0; i < n; ++i) rhs[i] = E[i] - UB[i];
    forward(rhs);
    ++sweeps;
    if (abs_resid) abs_resid[it] = residual();
  }
  return sweeps;
}

// conductionQ (conductionQ.f:35-98), FP64, operation by operation.
// Returns 0, or 3 when b(1) == 0 (tridag.for:10-12 stops there).
int oracle_conductionQ(int nz, const double *z, double dt, double Qn, double Qnp1, double *T,
                       const double *ti, const double *rhoc, double emiss, double *Tsurf,
                       double Fgeotherm, double *Fsurf) {
  const double sigSB = 5.6704e-8;
  std::vector<double> k(nz), alpha(nz, 0.0), gamma(nz, 0.0), a(nz), b(nz), c(nz), r(nz),
      gam(nz);
  for (int i = 0; i < nz; ++i) k[i] = (ti[i] * ti[i]) / rhoc[i];
  const double dz = 2. * z[0];
  const double beta = (dt / rhoc[0]) / (2. * (dz * dz));
  alpha[0] = beta * k[1];
  gamma[0] = beta * k[0];
  for (int i = 1; i < nz - 1; ++i) {
    const double buf = dt / (z[i + 1] - z[i - 1]);
    alpha[i] = (((2. * k[i + 1]) * buf) / (rhoc[i] + rhoc[i + 1])) / (z[i + 1] - z[i]);
    gamma[i] = (((2. * k[i]) * buf) / (rhoc[i] + rhoc[i + 1])) / (z[i] - z[i - 1]);
  }
  const double dl = z[nz - 1] - z[nz - 2];
  const double buf = dt / (dl * dl);
  gamma[nz - 1] = (k[nz - 1] * buf) / (2 * rhoc[nz - 1]);
  const double k1 = k[0] / dz;
  for (int i = 0; i < nz; ++i) {
    a[i] = -gamma[i];
    b[i] = (1. + alpha[i]) + gamma[i];
    c[i] = -alpha[i];
  }
  b[nz - 1] = 1. + gamma[nz - 1];
  const double Tr = *Tsurf;
  // Tr**3 and Tr**4 as repeated multiplication, left to right (the order the
  // compiled reference Fortran evaluates them in).
  const double Tr3 = (Tr * Tr) * Tr, Tr4 = Tr3 * Tr;
  const double arad = -(((3. * emiss) * sigSB) * Tr4);
  const double brad = ((2. * emiss) * sigSB) * Tr3;
  const double ann = (Qn - arad) / (k1 + brad);
  const double annp1 = (Qnp1 - arad) / (k1 + brad);
  const double bn = (k1 - brad) / (k1 + brad);
  b[0] = ((1. + alpha[0]) + gamma[0]) - gamma[0] * bn;
  r[0] = (gamma[0] * (annp1 + ann) + (((1. - alpha[0]) - gamma[0]) + gamma[0] * bn) * T[0]) +
         alpha[0] * T[1];
  for (int i = 1; i < nz - 1; ++i)
    r[i] = (gamma[i] * T[i - 1] + ((1. - alpha[i]) - gamma[i]) * T[i]) + alpha[i] * T[i + 1];
  r[nz - 1] = (gamma[nz - 1] * T[nz - 2] + (1. - gamma[nz - 1]) * T[nz - 1]) +
              ((dt / rhoc[nz - 1]) * Fgeotherm) / (z[nz - 1] - z[nz - 2]);
  // Thomas algorithm.
  if (b[0] == 0.) return 3;
  double bet = b[0];
  T[0] = r[0] / bet;
  for (int j = 1; j < nz; ++j) {
    gam[j] = c[j - 1] / bet;
    bet = b[j] - a[j] * gam[j];
    T[j] = (r[j] - a[j] * T[j - 1]) / bet;
  }
  for (int j = nz - 2; j >= 0; --j) T[j] = T[j] - gam[j + 1] * T[j + 1];
  *Tsurf = 0.5 * ((annp1 + bn * T[0]) + T[0]);
  *Fsurf = -((k[0] * (T[0] - *Tsurf)) / z[0]);
  return 0;
}

// thermal_model::step repeated nsteps times (thermal.hpp:65-84) for nf facets
// with per-facet layer grids: facet f uses depths z[goff[g] ..], g =
// grid[f] (grid null => 0), thermal inertia ti[...] * ti_scale[f].  State
// T: (nzmax+1) x nf column-major, row 0 = surface.  Q: nsteps x nf (row k is
// Q at step k+1), Qprev: nf (updated), Fsurf: nf (out).  Tsurf_series
// (nullable): nsteps x nf.
int oracle_thermal_run(int64_t nf, const int64_t *goff, const double *z, const double *ti,
                       const double *rhoc, const int32_t *grid, const double *ti_scale,
                       double emiss, int nzmax, double dt, int nsteps, const double *Q,
                       double *Qprev, double *T, double *Fsurf, double *Tsurf_series,
                       int nthreads) {
  std::atomic<int> status(0);
  parallel_for(nf, nthreads, [&](int64_t f) {
    const int g = grid ? grid[f] : 0;
    const int64_t o = goff[g];
    const int nz = (int)(goff[g + 1] - o);
    std::vector<double> tl(nz);
    const double s = ti_scale ? ti_scale[f] : 1.0;
    for (int i = 0; i < nz; ++i) tl[i] = ti_scale ? ti[o + i] * s : ti[o + i];
    double *Tf = T + f * (nzmax + 1);
    double qn = Qprev[f];
    for (int k = 0; k < nsteps; ++k) {
      const double qnp1 = Q[(int64_t)k * nf + f];
      const int st = oracle_conductionQ(nz, z + o, dt, qn, qnp1, Tf + 1, tl.data(), rhoc + o,
                                        emiss, Tf, 0.0, Fsurf + f);
      if (st) { status = st; return; }
      if (Tsurf_series) Tsurf_series[(int64_t)k * nf + f] = Tf[0];
      qn = qnp1;
    }
    Qprev[f] = qn;
  });
  return status.load();
}

// --------------------------------------------------------------------------
// Renderer ray-cast (illum_renderer_cli.cpp:174-261): facet hit by each pixel's
// ray, -1 for none, row-major height x width (row 0 at the top).
// mode 0: orthographic camera at azimuth az, elevation el (degrees) looking
// at the mesh centroid from 1.1 x the largest centroid distance; mode 1:
// equal-area cylindrical projection, az and el used as given (the reference
// converts degrees only in the orthographic branch).  Vector arithmetic is
// FP32 component-wise, scalars are FP64 rounded to FP32 where they meet a
// vector; |v| = sqrtf(x*x + y*y + z*z) summed left to right.
namespace {
struct V3 { float x, y, z; };
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float len(V3 a) { return (float)std::sqrt((double)((a.x * a.x + a.y * a.y) + a.z * a.z)); }
inline V3 unit(V3 a) { const float l = len(a); return {a.x / l, a.y / l, a.z / l}; }
inline V3 sph2xyz(double phi, double theta, double rad) {
  return {(float)((rad * std::cos(phi)) * std::sin(theta)),
          (float)((rad * std::sin(phi)) * std::sin(theta)), (float)(rad * std::cos(theta))};
}
}  // namespace

int oracle_render(const float *tri9, const float *cen, int64_t nf, int mode, int width,
                  int height, double az, double el, int nthreads, int32_t *facet) {
  const double PI = 3.141592653589793;
  V3 c{0.0f, 0.0f, 0.0f};
  for (int64_t i = 0; i < nf; ++i) c = add(c, V3{cen[3 * i], cen[3 * i + 1], cen[3 * i + 2]});
  c = V3{c.x / (float)nf, c.y / (float)nf, c.z / (float)nf};
  double r_model = 0.0;
  for (int64_t i = 0; i < nf; ++i)
    r_model = std::fmax(r_model, (double)len(sub(V3{cen[3 * i], cen[3 * i + 1], cen[3 * i + 2]}, c)));
  const double rad = 1.1 * r_model;
  BVH bvh;
  bvh.build(tri9, nf);
  V3 p_cam{}, n_up{}, n_right{}, n_eye{};
  double aspect = (double)width / height;
  if (mode == 0) {
    az *= PI / 180;
    el *= PI / 180;
    p_cam = sph2xyz(az, el, rad);
    const V3 u = unit(p_cam);
    n_eye = {-u.x, -u.y, -u.z};
    n_up = sph2xyz(az, el - PI / 2, 1.0);
    n_right = cross(n_eye, n_up);
    p_cam = add(p_cam, c);
  }
  parallel_for((int64_t)width * height, nthreads, [&](int64_t k) {
    const int i = (int)(k / width), j = (int)(k % width);
    V3 o, d;
    if (mode == 0) {
      double y_eye = 2 * (1. - (double)i / (height - 1)) - 1;
      y_eye *= rad;
      double x_eye = 2 * (double)j / (width - 1) - 1;
      x_eye *= aspect * rad;
      o = add(add(p_cam, mul((float)y_eye, n_up)), mul((float)x_eye, n_right));
      d = n_eye;
    } else {
      const double theta = std::asin(2 * (double)(i + 1) / (height + 1) - 1);
      const double phi = 2 * PI * (double)j / width;
      const V3 pc = sph2xyz(phi - az, theta - el, rad);
      const V3 u = unit(pc);
      d = {-u.x, -u.y, -u.z};
      o = add(pc, c);
    }
    const float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
    float tb;
    facet[k] = bvh.closest(setup(of, df), &tb, false);
  });
  return 0;
}

}  // extern "C"
