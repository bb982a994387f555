// Device ray/triangle predicate and stackless BVH traversal.
//
// Visibility is defined by this FP32 predicate (the reference delegates it to
// Fast-BVH's Tri::getIntersection, illum.cpp:73 and :301, whose source is not
// available).  It is the watertight test of Woop, Benthin & Wald, "Watertight
// Ray/Triangle Intersection", JCGT 2(1), 2013, two-sided, with the paper's
// FP64 recomputation of an edge function that evaluates to exactly zero.
// The reported distance is max(t, e), where e is the entry distance of the ray
// into the triangle's own bounding box.  Node boxes and the triangle's box use
// one slab function, s(x) = fma(x, inv, -o * inv) per axis (one correctly
// rounded multiply-add of the coordinate with per-ray constants), which is
// monotone in x; so every ancestor box has an entry distance <= the reported
// distance, and pruning boxes whose entry lies beyond the best hit never
// changes the closest hit: a traversal returns exactly what a brute-force
// loop over all triangles returns, whatever the tree.
// Closest hit = smallest reported distance, ties to the smaller facet index.
//
// All arithmetic is written out operation by operation; the library is built
// with -ffp-contract=off so no multiply-add is fused.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace illum {

struct DRay {
  float ox, oy, oz;   // origin
  float ix, iy, iz;   // per-axis inverse direction for box slabs (finite)
  float nx, ny, nz;   // -o * inv per axis (the slab function's offset)
  float Sx, Sy, Sz;   // shear constants
  int kx, ky, kz;     // axis permutation (kz = dominant axis)
};

// Wavefront ballot of a lane predicate, taken as a lane mask directly (no
// round trip of the predicate through an int).
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ float sel3(float x, float y, float z, int k) {
  return k == 0 ? x : (k == 1 ? y : z);
}

// Inverse of a direction component for the slab test.  A zero component is
// replaced by +-1e-30 so every slab product stays finite and monotone.
__device__ __forceinline__ float slab_inv(float d) {
  float dd = d == 0.0f ? (__float_as_uint(d) >> 31 ? -1e-30f : 1e-30f) : d;
  return 1.0f / dd;
}

__device__ __forceinline__ DRay make_ray(float ox, float oy, float oz, float dx, float dy,
                                         float dz) {
  DRay r;
  r.ox = ox; r.oy = oy; r.oz = oz;
  float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
  int kz = 0;
  float m = ax;
  if (ay > m) { kz = 1; m = ay; }
  if (az > m) { kz = 2; }
  int kx = kz + 1; if (kx == 3) kx = 0;
  int ky = kx + 1; if (ky == 3) ky = 0;
  float dkz = sel3(dx, dy, dz, kz);
  if (dkz < 0.0f) { int t = kx; kx = ky; ky = t; }
  r.kx = kx; r.ky = ky; r.kz = kz;
  r.Sx = sel3(dx, dy, dz, kx) / dkz;
  r.Sy = sel3(dx, dy, dz, ky) / dkz;
  r.Sz = 1.0f / dkz;
  r.ix = slab_inv(dx); r.iy = slab_inv(dy); r.iz = slab_inv(dz);
  r.nx = -(ox * r.ix); r.ny = -(oy * r.iy); r.nz = -(oz * r.iz);
  return r;
}

// Box acceptance: tn <= tf && tf > 0 && tn <= tbest, written as the single
// comparison max(tn, m) <= min(tf, tbest) with m the smallest positive float.
// Equivalent whenever tbest > 0 (every caller's best distance is a hit
// distance, a positive bound or +inf; a negative tbest rejects every box):
// tf >= m is tf > 0, and FP32 denormals are kept (the kernels' FP mode), so
// m and denormal tf compare as values.  One comparison gives a lane mask that
// a ballot takes as is.
__device__ __forceinline__ bool slab_accept(float tn, float tf, float tbest) {
  return fmaxf(tn, __int_as_float(1)) <= fminf(tf, tbest);
}

// Slab test against a node box: whether the box must be visited for a best
// distance tbest.
__device__ __forceinline__ bool box_visit(const DRay &r, float4 lo, float4 hi, float tbest) {
  float t0x = __builtin_fmaf(lo.x, r.ix, r.nx), t1x = __builtin_fmaf(hi.x, r.ix, r.nx);
  float t0y = __builtin_fmaf(lo.y, r.iy, r.ny), t1y = __builtin_fmaf(hi.y, r.iy, r.ny);
  float t0z = __builtin_fmaf(lo.z, r.iz, r.nz), t1z = __builtin_fmaf(hi.z, r.iz, r.nz);
  float tn = fmaxf(fminf(t0x, t1x), fmaxf(fminf(t0y, t1y), fminf(t0z, t1z)));
  float tf = fminf(fmaxf(t0x, t1x), fminf(fmaxf(t0y, t1y), fmaxf(t0z, t1z)));
  return slab_accept(tn, tf, tbest);
}

// Permutation class of a ray: f = 2 kz + (d_kz < 0), which fixes (kx, ky, kz)
// as make_ray chooses them.  PermOf<f> gives the axes at compile time.
template <int F>
struct PermOf {
  static constexpr int kz = F >> 1;
  static constexpr int kx = (F & 1) ? (kz + 2) % 3 : (kz + 1) % 3;
  static constexpr int ky = (F & 1) ? (kz + 1) % 3 : (kz + 2) % 3;
};

__device__ __forceinline__ int ray_perm(const DRay &r) {
  return 2 * r.kz + (r.kx == (r.kz == 0 ? 2 : r.kz - 1) ? 1 : 0);
}

// Watertight two-sided test.  On a hit (t > 0) writes the reported distance.
// PERM >= 0: the caller guarantees ray_perm(r) == PERM, so the axis
// selections are resolved at compile time (same operations, same result).
template <int PERM = -1>
__device__ __forceinline__ bool tri_hit(const DRay &r, float4 a, float4 b, float4 c,
                                        float *tout) {
  const int kx = PERM >= 0 ? PermOf<(PERM < 0 ? 0 : PERM)>::kx : r.kx;
  const int ky = PERM >= 0 ? PermOf<(PERM < 0 ? 0 : PERM)>::ky : r.ky;
  const int kz = PERM >= 0 ? PermOf<(PERM < 0 ? 0 : PERM)>::kz : r.kz;
  const float Ax0 = a.x - r.ox, Ay0 = a.y - r.oy, Az0 = a.z - r.oz;
  const float Bx0 = b.x - r.ox, By0 = b.y - r.oy, Bz0 = b.z - r.oz;
  const float Cx0 = c.x - r.ox, Cy0 = c.y - r.oy, Cz0 = c.z - r.oz;
  const float Akx = sel3(Ax0, Ay0, Az0, kx), Aky = sel3(Ax0, Ay0, Az0, ky),
              Akz = sel3(Ax0, Ay0, Az0, kz);
  const float Bkx = sel3(Bx0, By0, Bz0, kx), Bky = sel3(Bx0, By0, Bz0, ky),
              Bkz = sel3(Bx0, By0, Bz0, kz);
  const float Ckx = sel3(Cx0, Cy0, Cz0, kx), Cky = sel3(Cx0, Cy0, Cz0, ky),
              Ckz = sel3(Cx0, Cy0, Cz0, kz);
  const float Ax = Akx - r.Sx * Akz, Ay = Aky - r.Sy * Akz;
  const float Bx = Bkx - r.Sx * Bkz, By = Bky - r.Sy * Bkz;
  const float Cx = Ckx - r.Sx * Ckz, Cy = Cky - r.Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
    double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
    double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
    U = (float)(CxBy - CyBx);
    V = (float)(AxCy - AyCx);
    W = (float)(BxAy - ByAx);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = (U + V) + W;
  if (det == 0.0f) return false;
  const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
  const float T = (U * Az + V * Bz) + W * Cz;
  const float t = T / det;
  if (!(t > 0.0f)) return false;
  // Entry distance into the triangle's own bounding box (the slab function).
  const float ex = fminf(__builtin_fmaf(a.x, r.ix, r.nx),
                         fminf(__builtin_fmaf(b.x, r.ix, r.nx), __builtin_fmaf(c.x, r.ix, r.nx)));
  const float ey = fminf(__builtin_fmaf(a.y, r.iy, r.ny),
                         fminf(__builtin_fmaf(b.y, r.iy, r.ny), __builtin_fmaf(c.y, r.iy, r.ny)));
  const float ez = fminf(__builtin_fmaf(a.z, r.iz, r.nz),
                         fminf(__builtin_fmaf(b.z, r.iz, r.nz), __builtin_fmaf(c.z, r.iz, r.nz)));
  const float e = fmaxf(ex, fmaxf(ey, ez));
  *tout = fmaxf(t, e);
  return true;
}

// Whether tri_hit(r, a, b, c) reports a hit, for any-hit queries that need
// no distance: the same edge functions and sign tests, then t = T / det > 0
// decided without dividing.  t > 0 exactly when T and det have one sign and
// the quotient does not round to zero, i.e. |T| > 2^-150 |det|; that is
// tested as (|T| 2^75) 2^75 > |det|, whose two power-of-two scalings are
// exact (an overflow to +inf still compares greater).  FP32 denormals are
// kept (the kernels' FP mode).
__device__ __forceinline__ bool tri_any(const DRay &r, float4 a, float4 b, float4 c) {
  const float Ax0 = a.x - r.ox, Ay0 = a.y - r.oy, Az0 = a.z - r.oz;
  const float Bx0 = b.x - r.ox, By0 = b.y - r.oy, Bz0 = b.z - r.oz;
  const float Cx0 = c.x - r.ox, Cy0 = c.y - r.oy, Cz0 = c.z - r.oz;
  const float Akx = sel3(Ax0, Ay0, Az0, r.kx), Aky = sel3(Ax0, Ay0, Az0, r.ky),
              Akz = sel3(Ax0, Ay0, Az0, r.kz);
  const float Bkx = sel3(Bx0, By0, Bz0, r.kx), Bky = sel3(Bx0, By0, Bz0, r.ky),
              Bkz = sel3(Bx0, By0, Bz0, r.kz);
  const float Ckx = sel3(Cx0, Cy0, Cz0, r.kx), Cky = sel3(Cx0, Cy0, Cz0, r.ky),
              Ckz = sel3(Cx0, Cy0, Cz0, r.kz);
  const float Ax = Akx - r.Sx * Akz, Ay = Aky - r.Sy * Akz;
  const float Bx = Bkx - r.Sx * Bkz, By = Bky - r.Sy * Bkz;
  const float Cx = Ckx - r.Sx * Ckz, Cy = Cky - r.Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
    double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
    double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
    U = (float)(CxBy - CyBx);
    V = (float)(AxCy - AyCx);
    W = (float)(BxAy - ByAx);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  const float det = (U + V) + W;
  if (det == 0.0f) return false;
  const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
  const float T = (U * Az + V * Bz) + W * Cz;
  if (det > 0.0f ? !(T > 0.0f) : !(T < 0.0f)) return false;
  return (fabsf(T) * 0x1p75f) * 0x1p75f > fabsf(det);
}

// Closest hit over the threaded tree starting at node `root`; returns the
// original facet index or -1.  nbox / ntri count the box and triangle tests.
__device__ __forceinline__ int trace_closest(const float4 *__restrict__ nodes,
                                             const float4 *__restrict__ tris, const DRay &r,
                                             int root, float *tbest_out, unsigned &nbox,
                                             unsigned &ntri) {
  float tbest = __builtin_huge_valf();
  int best = -1;
  int node = root;
  while (node >= 0) {
    const float4 lo = nodes[2 * node], hi = nodes[2 * node + 1];
    const int miss = __float_as_int(lo.w), info = __float_as_int(hi.w);
    ++nbox;
    if (!box_visit(r, lo, hi, tbest)) { node = miss; continue; }
    if (info < 0) { node = node + 1; continue; }
    const int first = info & 0xFFFFFF, cnt = info >> 24;
    ntri += cnt;
    for (int k = 0; k < cnt; ++k) {
      const float4 a = tris[3 * (first + k)], b = tris[3 * (first + k) + 1],
                   c = tris[3 * (first + k) + 2];
      const int idx = __float_as_int(a.w);
      float t;
      if (tri_hit(r, a, b, c, &t) && (t < tbest || (t == tbest && idx < best))) {
        tbest = t;
        best = idx;
      }
    }
    node = miss;
  }
  *tbest_out = tbest;
  return best;
}

// Slab test returning the entry distance.
__device__ __forceinline__ bool box_enter(const DRay &r, float4 lo, float4 hi, float tbest,
                                          float *tn_out) {
  float t0x = __builtin_fmaf(lo.x, r.ix, r.nx), t1x = __builtin_fmaf(hi.x, r.ix, r.nx);
  float t0y = __builtin_fmaf(lo.y, r.iy, r.ny), t1y = __builtin_fmaf(hi.y, r.iy, r.ny);
  float t0z = __builtin_fmaf(lo.z, r.iz, r.nz), t1z = __builtin_fmaf(hi.z, r.iz, r.nz);
  float tn = fmaxf(fminf(t0x, t1x), fmaxf(fminf(t0y, t1y), fminf(t0z, t1z)));
  float tf = fminf(fmaxf(t0x, t1x), fminf(fmaxf(t0y, t1y), fmaxf(t0z, t1z)));
  *tn_out = tn;
  return slab_accept(tn, tf, tbest);
}

// One step of the near-first closest-hit walk over the two-wide layout: one
// inner node (descend into the nearer hit child, pushing the farther), or one
// leaf followed by a pop.  Returns true when the walk has finished.  `stack`
// points at this lane's first slot of an LDS stack laid out slot-major (entry
// k of the lane at stack[k * stride]) so lanes at equal depth touch
// consecutive words.  A push beyond `cap` entries sets sp = -1 and ends the
// walk (the caller then answers the query another way).  Boxes are pruned
// against the best distance so far; because every ancestor box enters no
// later than any hit inside it, a walk started with tbest = tinit ends with
// the argmin over all triangles with distance <= tinit.
__device__ __forceinline__ bool closest_wide_step(const float4 *__restrict__ wide,
                                                  const float4 *__restrict__ tris,
                                                  const DRay &r, int &cur, int &sp,
                                                  float &tbest, int &best, int *stack,
                                                  int stride, int cap, unsigned &nbox,
                                                  unsigned &ntri) {
  if (cur >= 0) {
    nbox += 2;
    const float4 a0 = wide[4 * cur], a1 = wide[4 * cur + 1];
    const float4 b0 = wide[4 * cur + 2], b1 = wide[4 * cur + 3];
    float tl, tr;
    const bool hl = box_enter(r, a0, a1, tbest, &tl);
    const bool hr = box_enter(r, b0, b1, tbest, &tr);
    const int cl = __float_as_int(a0.w), cr = __float_as_int(b0.w);
    if (hl && hr) {
      if (sp == cap) {
        sp = -1;
        return true;
      }
      const bool lnear = tl <= tr;
      stack[sp * stride] = lnear ? cr : cl;
      ++sp;
      cur = lnear ? cl : cr;
      return false;
    }
    if (hl) { cur = cl; return false; }
    if (hr) { cur = cr; return false; }
  } else {
    const int info = ~cur;
    const int first = info & 0xFFFFFF, cnt = info >> 24;
    ntri += cnt;
    for (int k = 0; k < cnt; ++k) {
      const float4 a = tris[3 * (first + k)], b = tris[3 * (first + k) + 1],
                   c = tris[3 * (first + k) + 2];
      const int idx = __float_as_int(a.w);
      float t;
      if (tri_hit(r, a, b, c, &t) && (t < tbest || (t == tbest && idx < best))) {
        tbest = t;
        best = idx;
      }
    }
  }
  if (sp == 0) return true;
  --sp;
  cur = stack[sp * stride];
  return false;
}

// Near-first closest hit over the two-wide tree rooted at inner node `root`
// (closest_wide_step to the end of the walk); returns the original facet
// index, -1 for no hit, or -2 when the walk needed more than `cap` stack
// entries.
__device__ __forceinline__ int trace_closest_wide(const float4 *__restrict__ wide,
                                                  const float4 *__restrict__ tris,
                                                  const DRay &r, int root, float tinit,
                                                  int *stack, int stride, int cap,
                                                  float *tbest_out, unsigned &nbox,
                                                  unsigned &ntri) {
  float tbest = tinit;
  int best = -1, sp = 0, cur = root;
  while (!closest_wide_step(wide, tris, r, cur, sp, tbest, best, stack, stride, cap, nbox,
                            ntri)) {
  }
  *tbest_out = tbest;
  return sp < 0 ? -2 : best;
}

// Wavefront-cooperative closest hit over the two-wide tree rooted at inner
// node `root`: all 64 lanes call it together (converged); lanes with
// act = false take no part.  The wavefront walks one node sequence, the union
// of its lanes' walks: a child is entered when any active lane's ray enters
// its box no later than that lane's best distance; with both children
// entered, the one more lanes enter first is visited first and the other
// pushed.  Node and triangle records are read at wavefront-uniform addresses
// (scalar loads, once per wavefront).  The stack is one VGPR across the
// lanes (entry k in lane k), so it needs no LDS; a walk pushes at most once
// per tree level, so depth + 1 <= 64 entries suffice (the caller checks).
// Each lane's result equals its own near-first walk's: every node the lane
// needs is visited (an entered child is visited by the wavefront), pruning
// uses the lane's own best distance, and a node a lane does not need only
// adds exact tests (its boxes lie inside the parent's, which the lane's ray
// does not enter before its best distance).
template <int PERM = -1>
__device__ __forceinline__ int trace_closest_packet(const float4 *__restrict__ wide,
                                                    const float4 *__restrict__ tris,
                                                    const DRay &r, bool act, int root,
                                                    float tinit, float *tbest_out,
                                                    unsigned &nbox, unsigned &ntri,
                                                    unsigned &wnode, unsigned &wtri) {
  const int lane = __lane_id();
  // an inactive lane's best distance is -inf: every box test rejects it
  float tbest = act ? tinit : -__builtin_huge_valf();
  int best = -1;
  int sp = 0;         // wavefront-uniform
  int cur = root;     // wavefront-uniform
  int stk = 0;        // lane k: stack entry k
  for (;;) {
    if (cur >= 0) {
      // the whole 64-byte node in one go, both boxes tested on every lane
      const float4 a0 = wide[4 * cur], a1 = wide[4 * cur + 1];
      const float4 b0 = wide[4 * cur + 2], b1 = wide[4 * cur + 3];
      float tl = 0.0f, tr = 0.0f;
      const bool hl = box_enter(r, a0, a1, tbest, &tl);
      const bool hr = box_enter(r, b0, b1, tbest, &tr);
      if (act) nbox += 2;
      ++wnode;
      const uint64_t ml = wave_ballot(hl), mr = wave_ballot(hr);
      const int cl = __float_as_int(a0.w), cr = __float_as_int(b0.w);
      if (ml && mr) {
        const uint64_t both = ml & mr;
        const uint64_t lfirst = wave_ballot(tl <= tr) & both;
        const bool lnear = 2 * __popcll(lfirst) >= __popcll(both);
        if (lane == sp) stk = lnear ? cr : cl;
        ++sp;
        cur = lnear ? cl : cr;
        continue;
      }
      if (ml) { cur = cl; continue; }
      if (mr) { cur = cr; continue; }
    } else {
      const int info = ~cur;
      const int first = info & 0xFFFFFF, cnt = info >> 24;
      if (act) ntri += cnt;
      wtri += cnt;
      for (int k = 0; k < cnt; ++k) {
        const float4 a = tris[3 * (first + k)], b = tris[3 * (first + k) + 1],
                     c = tris[3 * (first + k) + 2];
        const int idx = __float_as_int(a.w);
        float t;
        if (act && tri_hit<PERM>(r, a, b, c, &t) && (t < tbest || (t == tbest && idx < best))) {
          tbest = t;
          best = idx;
        }
      }
    }
    if (sp == 0) break;
    --sp;
    cur = __builtin_amdgcn_readlane(stk, sp);
  }
  *tbest_out = tbest;
  return best;
}

// Slab test of the four children of a four-wide node (SoA boxes); returns
// the entry distance, +inf for a child that is absent or need not be visited
// (same acceptance as box_enter against tbest).
__device__ __forceinline__ void box4_enter(const DRay &r, const float4 *__restrict__ nd,
                                           float tbest, float key[4], int ref[4]) {
  const float4 lx = nd[0], ly = nd[1], lz = nd[2], hx = nd[3], hy = nd[4], hz = nd[5];
  const float4 rf = nd[6];
  const int cnt = __float_as_int(nd[7].x);
  const float lox[4] = {lx.x, lx.y, lx.z, lx.w}, loy[4] = {ly.x, ly.y, ly.z, ly.w},
              loz[4] = {lz.x, lz.y, lz.z, lz.w}, hix[4] = {hx.x, hx.y, hx.z, hx.w},
              hiy[4] = {hy.x, hy.y, hy.z, hy.w}, hiz[4] = {hz.x, hz.y, hz.z, hz.w};
  const float rfa[4] = {rf.x, rf.y, rf.z, rf.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float t0x = __builtin_fmaf(lox[q], r.ix, r.nx), t1x = __builtin_fmaf(hix[q], r.ix, r.nx);
    const float t0y = __builtin_fmaf(loy[q], r.iy, r.ny), t1y = __builtin_fmaf(hiy[q], r.iy, r.ny);
    const float t0z = __builtin_fmaf(loz[q], r.iz, r.nz), t1z = __builtin_fmaf(hiz[q], r.iz, r.nz);
    const float tn = fmaxf(fminf(t0x, t1x), fmaxf(fminf(t0y, t1y), fminf(t0z, t1z)));
    const float tf = fminf(fmaxf(t0x, t1x), fminf(fmaxf(t0y, t1y), fmaxf(t0z, t1z)));
    const bool ok = q < cnt && slab_accept(tn, tf, tbest);
    key[q] = ok ? fminf(tn, 3.0e38f) : __builtin_huge_valf();
    ref[q] = __float_as_int(rfa[q]);
  }
}

__device__ __forceinline__ void cswap4(float *k, int *r, int a, int b) {
  if (k[b] < k[a]) {
    const float tk = k[a]; k[a] = k[b]; k[b] = tk;
    const int tr = r[a]; r[a] = r[b]; r[b] = tr;
  }
}

// Wavefront-cooperative closest hit over the four-wide tree rooted at `root`
// (same contract and exactness argument as trace_closest_packet).  A child is
// entered when any active lane enters it; entered children are visited in
// order of the entry distance seen by the first lane that enters each (the
// lanes share an origin and aim at neighbouring points, so one lane's order
// is representative), the nearest next and the others pushed farthest first.
// A walk pushes at most three entries per level, so 3 * depth4 + 1 <= 64
// entries suffice (the caller checks).
template <int PERM = -1>
__device__ __forceinline__ int trace_closest_packet4(const float4 *__restrict__ w4,
                                                     const float4 *__restrict__ tris,
                                                     const DRay &r, bool act, int root,
                                                     float tinit, float *tbest_out,
                                                     unsigned &nbox, unsigned &ntri,
                                                     unsigned &wnode, unsigned &wtri) {
  const int lane = __lane_id();
  // an inactive lane's best distance is -inf: every box test rejects it
  float tbest = act ? tinit : -__builtin_huge_valf();
  int best = -1;
  int sp = 0;
  int cur = root;
  int stk = 0;
  for (;;) {
    if (cur >= 0) {
      float key[4];
      int ref[4];
      box4_enter(r, w4 + 8 * cur, tbest, key, ref);
      if (act) nbox += 4;
      ++wnode;
      float k[4];
      int n = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint64_t m = wave_ballot(key[q] < __builtin_huge_valf());
        // representative entry distance: the first entering lane's
        k[q] = m ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(key[q]),
                                                              (int)__builtin_ctzll(m)))
                 : __builtin_huge_valf();
        n += m ? 1 : 0;
      }
      if (n > 0) {
        cswap4(k, ref, 0, 1);
        cswap4(k, ref, 2, 3);
        cswap4(k, ref, 0, 2);
        cswap4(k, ref, 1, 3);
        cswap4(k, ref, 1, 2);
        if (n > 3) { if (lane == sp) stk = ref[3]; ++sp; }
        if (n > 2) { if (lane == sp) stk = ref[2]; ++sp; }
        if (n > 1) { if (lane == sp) stk = ref[1]; ++sp; }
        cur = ref[0];
        continue;
      }
    } else {
      const int info = ~cur;
      const int first = info & 0xFFFFFF, cnt = info >> 24;
      if (act) ntri += cnt;
      wtri += cnt;
      for (int q = 0; q < cnt; ++q) {
        const float4 a = tris[3 * (first + q)], b = tris[3 * (first + q) + 1],
                     c = tris[3 * (first + q) + 2];
        const int idx = __float_as_int(a.w);
        float t;
        if (act && tri_hit<PERM>(r, a, b, c, &t) && (t < tbest || (t == tbest && idx < best))) {
          tbest = t;
          best = idx;
        }
      }
    }
    if (sp == 0) break;
    --sp;
    cur = __builtin_amdgcn_readlane(stk, sp);
  }
  *tbest_out = tbest;
  return best;
}

// Near-first closest hit over the four-wide tree rooted at `root`: the
// entered children are sorted by entry distance (5-comparator network), the
// nearest is visited next and the others pushed farthest first.  Returns the
// facet index, -1 for no hit, or -2 when a push would exceed `cap` entries.
// Exact for the same reason as the two-wide walk (ancestor boxes enter no
// later than any hit inside them; ties go to the smaller index).
__device__ __forceinline__ int trace_closest_wide4(const float4 *__restrict__ w4,
                                                   const float4 *__restrict__ tris,
                                                   const DRay &r, int root, float tinit,
                                                   int *stack, int stride, int cap,
                                                   float *tbest_out, unsigned &nbox,
                                                   unsigned &ntri) {
  float tbest = tinit;
  int best = -1, sp = 0, cur = root;
  for (;;) {
    if (cur >= 0) {
      float key[4];
      int ref[4];
      box4_enter(r, w4 + 8 * cur, tbest, key, ref);
      nbox += 4;
      cswap4(key, ref, 0, 1);
      cswap4(key, ref, 2, 3);
      cswap4(key, ref, 0, 2);
      cswap4(key, ref, 1, 3);
      cswap4(key, ref, 1, 2);
      const int nh = (key[0] < __builtin_huge_valf()) + (key[1] < __builtin_huge_valf()) +
                     (key[2] < __builtin_huge_valf()) + (key[3] < __builtin_huge_valf());
      if (nh > 0) {
        if (sp + nh - 1 > cap) {
          sp = -1;
          break;
        }
        if (nh > 3) stack[(sp++) * stride] = ref[3];
        if (nh > 2) stack[(sp++) * stride] = ref[2];
        if (nh > 1) stack[(sp++) * stride] = ref[1];
        cur = ref[0];
        continue;
      }
    } else {
      const int info = ~cur;
      const int first = info & 0xFFFFFF, cnt = info >> 24;
      ntri += cnt;
      for (int k = 0; k < cnt; ++k) {
        const float4 a = tris[3 * (first + k)], b = tris[3 * (first + k) + 1],
                     c = tris[3 * (first + k) + 2];
        const int idx = __float_as_int(a.w);
        float t;
        if (tri_hit(r, a, b, c, &t) && (t < tbest || (t == tbest && idx < best))) {
          tbest = t;
          best = idx;
        }
      }
    }
    if (sp == 0) break;
    --sp;
    cur = stack[sp * stride];
  }
  *tbest_out = tbest;
  return sp < 0 ? -2 : best;
}

// Any hit (t > 0) over the four-wide tree rooted at `root`: returns 1 on a
// hit, 0 for none, -2 when a push would exceed `cap` stack entries.  On a
// hit, *hit_tri (nullable) receives the hit triangle's leaf-order index.
__device__ __forceinline__ int trace_any_wide4(const float4 *__restrict__ w4,
                                               const float4 *__restrict__ tris, const DRay &r,
                                               int root, int *stack, int stride, int cap,
                                               int *hit_tri = nullptr) {
  int sp = 0, cur = root;
  for (;;) {
    if (cur >= 0) {
      float key[4];
      int ref[4];
      box4_enter(r, w4 + 8 * cur, __builtin_huge_valf(), key, ref);
      cswap4(key, ref, 0, 1);
      cswap4(key, ref, 2, 3);
      cswap4(key, ref, 0, 2);
      cswap4(key, ref, 1, 3);
      cswap4(key, ref, 1, 2);
      const int nh = (key[0] < __builtin_huge_valf()) + (key[1] < __builtin_huge_valf()) +
                     (key[2] < __builtin_huge_valf()) + (key[3] < __builtin_huge_valf());
      if (nh > 0) {
        if (sp + nh - 1 > cap) return -2;
        if (nh > 3) stack[(sp++) * stride] = ref[3];
        if (nh > 2) stack[(sp++) * stride] = ref[2];
        if (nh > 1) stack[(sp++) * stride] = ref[1];
        cur = ref[0];
        continue;
      }
    } else {
      const int info = ~cur;
      const int first = info & 0xFFFFFF, cnt = info >> 24;
      for (int k = 0; k < cnt; ++k) {
        float t;
        if (tri_any(r, tris[3 * (first + k)], tris[3 * (first + k) + 1],
                    tris[3 * (first + k) + 2])) {
          if (hit_tri) *hit_tri = first + k;
          return 1;
        }
      }
    }
    if (sp == 0) return 0;
    --sp;
    cur = stack[sp * stride];
  }
}

// Any hit (t > 0) over the two-wide tree rooted at `root`, nearer child
// first (stack as in trace_closest_wide; the tree depth must fit it).  On a
// hit, *hit_tri (nullable) receives the hit triangle's leaf-order index.
__device__ __forceinline__ bool trace_any_wide(const float4 *__restrict__ wide,
                                               const float4 *__restrict__ tris, const DRay &r,
                                               int root, int *stack, int stride,
                                               int *hit_tri = nullptr) {
  const float inf = __builtin_huge_valf();
  int sp = 0;
  int cur = root;
  for (;;) {
    if (cur >= 0) {
      const float4 a0 = wide[4 * cur], a1 = wide[4 * cur + 1];
      const float4 b0 = wide[4 * cur + 2], b1 = wide[4 * cur + 3];
      float tl, tr;
      const bool hl = box_enter(r, a0, a1, inf, &tl);
      const bool hr = box_enter(r, b0, b1, inf, &tr);
      const int cl = __float_as_int(a0.w), cr = __float_as_int(b0.w);
      if (hl && hr) {
        const bool lnear = tl <= tr;
        stack[sp * stride] = lnear ? cr : cl;
        ++sp;
        cur = lnear ? cl : cr;
        continue;
      }
      if (hl) { cur = cl; continue; }
      if (hr) { cur = cr; continue; }
    } else {
      const int info = ~cur;
      const int first = info & 0xFFFFFF, cnt = info >> 24;
      for (int k = 0; k < cnt; ++k) {
        float t;
        if (tri_any(r, tris[3 * (first + k)], tris[3 * (first + k) + 1],
                    tris[3 * (first + k) + 2])) {
          if (hit_tri) *hit_tri = first + k;
          return true;
        }
      }
    }
    if (sp == 0) break;
    --sp;
    cur = stack[sp * stride];
  }
  return false;
}

// Wavefront-cooperative any hit (t > 0) over the two-wide tree rooted at
// `root`, with the contract of trace_closest_packet: all 64 lanes call it
// together, lanes with act = false take no part.  A lane drops out of the
// walk once it has a hit; the walk ends when no lane is left searching or the
// tree is exhausted.  A child is entered when any searching lane's ray enters
// its box; with both entered, the one more lanes enter first goes first.
// Exact: a lane's answer is whether any triangle is hit, every node whose box
// its ray enters is visited until it has one, and extra tests only add exact
// hits.
__device__ __forceinline__ bool trace_any_packet(const float4 *__restrict__ wide,
                                                 const float4 *__restrict__ tris,
                                                 const DRay &r, bool act, int root) {
  const int lane = __lane_id();
  const float inf = __builtin_huge_valf();
  bool hit = false;
  int sp = 0;
  int cur = root;
  int stk = 0;
  for (;;) {
    const bool live = act && !hit;
    if (!wave_ballot(live)) break;
    if (cur >= 0) {
      const float4 a0 = wide[4 * cur], a1 = wide[4 * cur + 1];
      const float4 b0 = wide[4 * cur + 2], b1 = wide[4 * cur + 3];
      float tl, tr;
      // a lane not searching tests against -inf: every box rejects it
      const float tb = live ? inf : -inf;
      const bool hl = box_enter(r, a0, a1, tb, &tl);
      const bool hr = box_enter(r, b0, b1, tb, &tr);
      const uint64_t ml = wave_ballot(hl), mr = wave_ballot(hr);
      const int cl = __float_as_int(a0.w), cr = __float_as_int(b0.w);
      if (ml && mr) {
        const uint64_t both = ml & mr;
        const uint64_t lfirst = wave_ballot(tl <= tr) & both;
        const bool lnear = 2 * __popcll(lfirst) >= __popcll(both);
        if (lane == sp) stk = lnear ? cr : cl;
        ++sp;
        cur = lnear ? cl : cr;
        continue;
      }
      if (ml) { cur = cl; continue; }
      if (mr) { cur = cr; continue; }
    } else {
      const int info = ~cur;
      const int first = info & 0xFFFFFF, cnt = info >> 24;
      for (int k = 0; k < cnt; ++k) {
        float t;
        const bool h = tri_any(r, tris[3 * (first + k)], tris[3 * (first + k) + 1],
                               tris[3 * (first + k) + 2]);
        if (live && h) hit = true;
      }
    }
    if (sp == 0) break;
    --sp;
    cur = __builtin_amdgcn_readlane(stk, sp);
  }
  return hit;
}

// Any hit (t > 0) over the threaded tree starting at node `root`.  On a
// hit, *hit_tri (nullable) receives the hit triangle's leaf-order index.
__device__ __forceinline__ bool trace_any(const float4 *__restrict__ nodes,
                                          const float4 *__restrict__ tris, const DRay &r,
                                          int root, int *hit_tri = nullptr) {
  const float inf = __builtin_huge_valf();
  int node = root;
  while (node >= 0) {
    const float4 lo = nodes[2 * node], hi = nodes[2 * node + 1];
    const int miss = __float_as_int(lo.w), info = __float_as_int(hi.w);
    if (!box_visit(r, lo, hi, inf)) { node = miss; continue; }
    if (info < 0) { node = node + 1; continue; }
    const int first = info & 0xFFFFFF, cnt = info >> 24;
    for (int k = 0; k < cnt; ++k) {
      float t;
      if (tri_any(r, tris[3 * (first + k)], tris[3 * (first + k) + 1],
                  tris[3 * (first + k) + 2])) {
        if (hit_tri) *hit_tri = first + k;
        return true;
      }
    }
    node = miss;
  }
  return false;
}

}  // namespace illum
