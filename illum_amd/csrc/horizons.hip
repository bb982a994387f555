// Horizon maps (illum.cpp:274-363) and direct insolation (illum.cpp:382-483).
//
// Horizons: one lane per (facet, azimuth); a wavefront covers 64 consecutive
// azimuths of one facet, so its rays share an origin and sweep neighbouring
// directions (coherent traversal).  Each lane runs the reference bisection on
// zenith angle (illum.cpp:304-316): start at pi/2, a hit moves the upper bound
// down, a miss moves the lower bound up, stop when the midpoint moves by at
// most theta_eps.  The midpoints form a fixed binary tree; it is tabulated on
// the host (theta, sin theta, cos theta, stop flag per tree node) so the
// kernel evaluates exactly the reference's directions with the host's libm
// values and no device transcendental.  Only the boolean hit matters, so rays
// use the any-hit traversal.
#include "hip_util.hpp"
#include "ray.hpp"

#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

namespace illum {

// LAYOUT 4: near-first any-hit over the four-wide layout with a per-lane
// stack of `scap` entries in dynamic LDS (a walk needing more falls back to
// the threaded walk, same answer); LAYOUT 2: the two-wide layout with a
// (depth + 1)-entry stack; LAYOUT 0: the stackless threaded walk.
//
// Hit cache (CACHE): a lane keeps the triangle its last hit came from and
// tests it first; if the ray hits it, the answer (some triangle of the
// searched trees is hit) is known without a walk.  Exact: the cached
// triangle is one of the searched trees' triangles and the test is the
// walk's own predicate.  The bisection's rays after a hit rise towards the
// same ridge, and its last steps differ by under a degree, so most of them
// hit the cached triangle again.
template <int LAYOUT, bool CACHE>
__global__ __launch_bounds__(256) void horizons_kernel(
    int64_t nf, int nphi, const float4 *__restrict__ pr4, const float4 *__restrict__ n4,
    const double *__restrict__ frame, const float4 *__restrict__ nodes,
    const float4 *__restrict__ wide, const float4 *__restrict__ wide4,
    const float4 *__restrict__ tris, float offset,
    const double *__restrict__ cphi, const double *__restrict__ sphi,
    const double *__restrict__ th, int tree_nodes, double *__restrict__ out,
    const int32_t *__restrict__ roots, int nroots, const int32_t *__restrict__ body_of,
    int scap) {
  extern __shared__ int hstack[];
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nf * nphi) return;
  const int64_t j = idx / nphi;
  const int k = (int)(idx - j * nphi);
  const float4 p = pr4[j], n = n4[j];
  const float ox = p.x + offset * n.x, oy = p.y + offset * n.y, oz = p.z + offset * n.z;
  const double *F = frame + 9 * j;
  const double F00 = F[0], F10 = F[1], F20 = F[2];  // column b
  const double F01 = F[3], F11 = F[4], F21 = F[5];  // column t
  const double F02 = F[6], F12 = F[7], F22 = F[8];  // column n
  const double cp = cphi[k], sp = sphi[k];
  int m = 1;
  int last = -1;  // leaf-order index of the last hit triangle (CACHE)
  for (;;) {
    const double st = th[4 * m + 1], ct = th[4 * m + 2];
    const double vx = cp * st, vy = sp * st, vz = ct;
    const float dx = (float)((F00 * vx + F01 * vy) + F02 * vz);
    const float dy = (float)((F10 * vx + F11 * vy) + F12 * vz);
    const float dz = (float)((F20 * vx + F21 * vy) + F22 * vz);
    const DRay r = make_ray(ox, oy, oz, dx, dy, dz);
    // Rays see the facet's own body (body_of set), else every tree in turn
    // (one body, or an occluder mesh).
    const int b0 = body_of ? body_of[j] : 0, b1 = body_of ? b0 + 1 : nroots;
    bool hit = false;
    int *ht = CACHE ? &last : nullptr;
    if (CACHE && last >= 0)
      hit = tri_any(r, tris[3 * last], tris[3 * last + 1], tris[3 * last + 2]);
    for (int b = b0; b < b1 && !hit; ++b) {
      if (LAYOUT == 4) {
        const int a = trace_any_wide4(wide4, tris, r, roots[4 * b + 2], hstack + threadIdx.x,
                                      256, scap, ht);
        hit = a == -2 ? trace_any(nodes, tris, r, roots[4 * b], ht) : a == 1;
      } else if (LAYOUT == 2) {
        hit = trace_any_wide(wide, tris, r, roots[4 * b + 1], hstack + threadIdx.x, 256, ht);
      } else {
        hit = trace_any(nodes, tris, r, roots[4 * b], ht);
      }
    }
    m = 2 * m + (hit ? 0 : 1);
    if (m >= tree_nodes) { out[idx] = __builtin_nan(""); return; }
    if (th[4 * m + 3] != 0.0) break;
  }
  out[idx] = th[4 * m];
}

// Wavefront packet variant (trace_any_packet).  A wavefront's lanes are 64
// consecutive (facet, azimuth) pairs, mostly of one facet, so their rays
// share an origin.  The lanes step their bisections together; a lane whose
// search has ended, or that is past the end, takes no part.  The wavefront
// walks the trees of the union of its lanes' bodies in turn, each lane
// testing only its own.  Per lane this is the same sequence of queries as
// horizons_kernel, so the same result.
__global__ __launch_bounds__(256) void horizons_packet_kernel(
    int64_t nf, int nphi, const float4 *__restrict__ pr4, const float4 *__restrict__ n4,
    const double *__restrict__ frame, const float4 *__restrict__ wide,
    const float4 *__restrict__ tris, float offset, const double *__restrict__ cphi,
    const double *__restrict__ sphi, const double *__restrict__ th, int tree_nodes,
    double *__restrict__ out, const int32_t *__restrict__ roots, int nroots,
    const int32_t *__restrict__ body_of) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = idx < nf * nphi;
  const int64_t j = valid ? idx / nphi : 0;
  const int k = (int)(idx - j * nphi);
  const float4 p = pr4[j], n = n4[j];
  const float ox = p.x + offset * n.x, oy = p.y + offset * n.y, oz = p.z + offset * n.z;
  const double *F = frame + 9 * j;
  const double F00 = F[0], F10 = F[1], F20 = F[2];  // column b
  const double F01 = F[3], F11 = F[4], F21 = F[5];  // column t
  const double F02 = F[6], F12 = F[7], F22 = F[8];  // column n
  const double cp = valid ? cphi[k] : 1.0, sp = valid ? sphi[k] : 0.0;
  // the lane's trees [b0, b1); the wavefront's union [ulo, uhi)
  const int b0 = body_of ? body_of[j] : 0, b1 = body_of ? b0 + 1 : nroots;
  int ulo = b0, uhi = b1;
  if (body_of) {
    for (int o = 32; o > 0; o >>= 1) {
      ulo = min(ulo, __shfl_xor(ulo, o));
      uhi = max(uhi, __shfl_xor(uhi, o));
    }
  }
  int m = 1;
  bool done = !valid;
  double res = 0.0;
  while (__ballot(!done)) {
    const double st = th[4 * m + 1], ct = th[4 * m + 2];
    const double vx = cp * st, vy = sp * st, vz = ct;
    const float dx = (float)((F00 * vx + F01 * vy) + F02 * vz);
    const float dy = (float)((F10 * vx + F11 * vy) + F12 * vz);
    const float dz = (float)((F20 * vx + F21 * vy) + F22 * vz);
    const DRay r = make_ray(ox, oy, oz, dx, dy, dz);
    bool hit = false;
    for (int b = ulo; b < uhi; ++b) {
      const bool act = !done && !hit && b >= b0 && b < b1;
      if (!__ballot(act)) continue;
      const bool h = trace_any_packet(wide, tris, r, act, roots[4 * b + 1]);
      if (act && h) hit = true;
    }
    if (!done) {
      m = 2 * m + (hit ? 0 : 1);
      if (m >= tree_nodes) {
        res = __builtin_nan("");
        done = true;
      } else if (th[4 * m + 3] != 0.0) {
        res = th[4 * m];
        done = true;
      }
    }
  }
  if (valid) out[idx] = res;
}

// Tabulates the bisection tree: node m (root 1) holds theta (the midpoint
// tested at m), sin, cos and whether the search stops on arriving at m.
static std::vector<double> bisection_table(double theta_eps, int *tree_nodes) {
  const int max_depth = 22;
  std::vector<double> lo(2, 0.0), hi(2, M_PI), th(8, 0.0);
  // node 1
  std::vector<double> tab(4 * 2, 0.0);
  tab[4 * 1 + 0] = M_PI / 2;
  int depth = 0;
  std::vector<int> frontier = {1};
  std::vector<double> flo = {0.0}, fhi = {M_PI};
  while (!frontier.empty()) {
    if (depth >= max_depth) throw Error(ILLUM_ERR_ARG, "theta_eps too small (bisection deeper than 22 levels)");
    std::vector<int> nf;
    std::vector<double> nlo, nhi;
    const size_t need = (size_t)4 << (depth + 2);
    if (tab.size() < need) tab.resize(need, 0.0);
    for (size_t q = 0; q < frontier.size(); ++q) {
      const int m = frontier[q];
      const double mid = tab[4 * m];
      for (int c = 0; c < 2; ++c) {  // c = 0: hit (hi = mid), c = 1: miss (lo = mid)
        const double l = c == 0 ? flo[q] : mid, h = c == 0 ? mid : fhi[q];
        const double mid2 = (l + h) / 2;
        const int child = 2 * m + c;
        tab[4 * child] = mid2;
        const bool stop = !(std::fabs(mid2 - mid) > theta_eps);
        tab[4 * child + 3] = stop ? 1.0 : 0.0;
        if (!stop) { nf.push_back(child); nlo.push_back(l); nhi.push_back(h); }
      }
    }
    frontier.swap(nf); flo.swap(nlo); fhi.swap(nhi);
    ++depth;
  }
  const int nodes = (int)(tab.size() / 4);
  for (int m = 1; m < nodes; ++m) {
    tab[4 * m + 1] = std::sin(tab[4 * m]);
    tab[4 * m + 2] = std::cos(tab[4 * m]);
  }
  *tree_nodes = nodes;
  return tab;
}

void horizons(illum_bvh *bvh, illum_bvh *occl, int nphi, double theta_eps, double offset,
              double *out, int flags) {
  illum_mesh *mesh = bvh->mesh;
  illum_ctx *ctx = mesh->ctx;
  hipStream_t s = stream_of(ctx);
  if (nphi < 2) throw Error(ILLUM_ERR_ARG, "nphi must be >= 2");
  if (!(theta_eps > 0)) throw Error(ILLUM_ERR_ARG, "theta_eps must be positive");
  if (!out) throw Error(ILLUM_ERR_ARG, "null output");
  illum_bvh *target = occl ? occl : bvh;
  check_bvh_bodies(bvh);
  check_bvh_bodies(target);
  const int64_t nf = mesh->g.nf;
  int tree_nodes = 0;
  std::vector<double> tab = bisection_table(theta_eps, &tree_nodes);
  // phis = linspace(0, 2 pi, nphi) with the endpoint exact (illum.cpp:329).
  std::vector<double> cs(nphi), sn(nphi);
  const double dphi = 2 * M_PI / (nphi - 1);
  for (int k = 0; k < nphi; ++k) {
    const double ph = k == nphi - 1 ? 2 * M_PI : k * dphi;
    cs[k] = std::cos(ph);
    sn[k] = std::sin(ph);
  }
  DBuf<double> d_tab(tab.size()), d_c(nphi), d_s(nphi);
  ILLUM_HIP(hipMemcpyAsync(d_tab.p, tab.data(), 8 * tab.size(), hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(d_c.p, cs.data(), 8 * nphi, hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(d_s.p, sn.data(), 8 * nphi, hipMemcpyHostToDevice, s));
  const int64_t total = nf * nphi;
  DBuf<double> d_out;
  double *dst = out;
  if (!(flags & ILLUM_OUT_ON_DEVICE)) { d_out.reset(total); dst = d_out.p; }
  {
    StageTimer st(ctx, "horizons");
    const char *e = getenv("ILLUM_HZ_TRAVERSAL");
    // ILLUM_HZ_TRAVERSAL: "threaded", "wide4", "packet" (wavefront walk over
    // the two-wide tree), or the per-lane two-wide walk (default; faster than
    // four-wide for these short bisection rays)
    const std::string ev = e ? e : "";
    const int layout = ev == "threaded" ? 0
                       : ev == "wide4" ? (target->d_wide4 ? 4 : 0)
                       : ev == "packet" ? (target->d_wide && target->h.depth + 1 <= 64 ? 3 : 0)
                                        : (target->d_wide && target->h.depth <= 32 ? 2 : 0);
    const int scap = layout == 4 ? std::min(3 * target->h.depth4 + 1, 16)
                     : layout == 2 ? target->h.depth + 1 : 0;
    const size_t lds = (size_t)scap * 256 * 4;
    if (layout == 3) {
      horizons_packet_kernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(
          nf, nphi, (const float4 *)mesh->d_pr4, (const float4 *)mesh->d_n4, mesh->d_frame,
          (const float4 *)target->d_wide, (const float4 *)target->d_tris, (float)offset, d_c.p,
          d_s.p, d_tab.p, tree_nodes, dst, target->d_roots, target->nbody,
          occl ? nullptr : mesh->d_body);
    } else {
      // ILLUM_HZ_CACHE=0 turns the hit cache off (same horizons, for comparison)
      const char *ce = getenv("ILLUM_HZ_CACHE");
      const bool cache = !(ce && std::string(ce) == "0");
      auto kern = layout == 4 ? (cache ? horizons_kernel<4, true> : horizons_kernel<4, false>)
                  : layout == 2 ? (cache ? horizons_kernel<2, true> : horizons_kernel<2, false>)
                                : (cache ? horizons_kernel<0, true> : horizons_kernel<0, false>);
      kern<<<(unsigned)((total + 255) / 256), 256, lds, s>>>(
          nf, nphi, (const float4 *)mesh->d_pr4, (const float4 *)mesh->d_n4, mesh->d_frame,
          (const float4 *)target->d_nodes, (const float4 *)target->d_wide,
          (const float4 *)target->d_wide4, (const float4 *)target->d_tris, (float)offset, d_c.p,
          d_s.p, d_tab.p, tree_nodes, dst, target->d_roots, target->nbody,
          occl ? nullptr : mesh->d_body, scap);
    }
    ILLUM_LAUNCH_CHECK();
  }
  if (!(flags & ILLUM_OUT_ON_DEVICE))
    ILLUM_HIP(hipMemcpyAsync(out, dst, 8 * total, hipMemcpyDeviceToHost, s));
  sync(ctx);
}

// ILLUM_DIRECT_EXACT=1 evaluates every disk sample (no whole-disk decision);
// results are identical either way, the switch exists for comparison.
static bool direct_exact_only() {
  const char *e = getenv("ILLUM_DIRECT_EXACT");
  return e && std::string(e) == "1";
}

// Lanes per (facet, sun) pair: enough for about 2^20 lanes (4096 workgroups),
// at most 16 (0, -4, -8, -16: the wavefront-cooperative kernel with 1, 4, 8
// or 16 lanes per pair; DIRECT_SPLIT: the two-pass shape with a queue of open
// pairs).  ILLUM_DIRECT_LANES overrides.
constexpr int DIRECT_SPLIT = 64;
static int direct_lanes(int64_t pairs) {
  if (const char *e = getenv("ILLUM_DIRECT_LANES")) {
    const int v = atoi(e);
    if (v == -16 || v == -8 || v == -4 || v == 0 || v == 1 || v == 2 || v == 4 || v == 8 ||
        v == 16 || v == DIRECT_SPLIT)
      return v;
  }
  int L = 1;
  while (L < 16 && pairs * L < ((int64_t)1 << 20)) L *= 2;
  return L;
}

// ---------------------------------------------------------------------------
// Direct insolation.  Two launch shapes with the same per-pair and
// per-sample operations (so the same E):
//   direct_kernel<L>: L adjacent lanes per (facet, sun direction), each
//     taking every L-th disk sample, their counts summed by lane shuffles (an
//     integer sum, so E does not depend on L).
//   direct_wave_kernel<P>: P lanes per pair for the set-up and the whole-disk
//     decision (P > 1 only adds wavefronts, to hide latency when the pairs
//     are few); the pairs of a wavefront that the decision leaves open are
//     then sampled one after the other by all 64 lanes.
// E is nf x S column-major so consecutive pairs write consecutive facets.
constexpr double ONE_AU_KM = 149.5978707e6;   // constants.hpp:7
constexpr double SUN_RADIUS_KM = 695508;      // constants.hpp:8

// One (facet, sun) pair: the facet's point and the sun disk's frame (N
// toward the sun centre, T and B spanning the disk).
struct DirectPair {
  double px, py, pz, sx, sy, sz;
  double Nx, Ny, Nz, Tx, Ty, Tz, Bx, By, Bz, nN;
  int64_t f;
  int s;
};

// Set-up of pair idx < nf S (facet idx mod nf, sun column idx / nf).
__device__ __forceinline__ DirectPair direct_setup(int64_t idx, int64_t nf, int S,
                                                   const float4 *__restrict__ pr4,
                                                   const double *__restrict__ sun,
                                                   const int32_t *__restrict__ body_of,
                                                   int per_body, const int *__restrict__ step) {
  DirectPair q;
  q.f = idx % nf;
  q.s = (int)(idx / nf);
  const float4 pf = pr4[q.f];
  q.px = pf.x; q.py = pf.y; q.pz = pf.z;
  // per_body: body b's S suns are columns b S .. b S + S - 1 of sun; with a
  // device step counter, step k reads the k-th such block of suns
  int64_t sc = per_body ? (int64_t)(body_of ? body_of[q.f] : 0) * S + q.s : q.s;
  if (step) sc += (int64_t)(*step) * S;
  q.sx = sun[3 * sc]; q.sy = sun[3 * sc + 1]; q.sz = sun[3 * sc + 2];
  double Nx = q.sx - q.px, Ny = q.sy - q.py, Nz = q.sz - q.pz;
  q.nN = sqrt((Nx * Nx + Ny * Ny) + Nz * Nz);
  Nx = Nx / q.nN; Ny = Ny / q.nN; Nz = Nz / q.nN;
  q.Nx = Nx; q.Ny = Ny; q.Nz = Nz;
  // deterministic tangent: project the least-aligned coordinate axis
  const double ax = fabs(Nx), ay = fabs(Ny), az = fabs(Nz);
  int e = 0;
  double mn = ax;
  if (ay < mn) { e = 1; mn = ay; }
  if (az < mn) { e = 2; }
  const double Ne = e == 0 ? Nx : (e == 1 ? Ny : Nz);
  double Tx = (e == 0 ? 1.0 : 0.0) - Nx * Ne;
  double Ty = (e == 1 ? 1.0 : 0.0) - Ny * Ne;
  double Tz = (e == 2 ? 1.0 : 0.0) - Nz * Ne;
  const double nT = sqrt((Tx * Tx + Ty * Ty) + Tz * Tz);
  Tx = Tx / nT; Ty = Ty / nT; Tz = Tz / nT;
  q.Tx = Tx; q.Ty = Ty; q.Tz = Tz;
  q.Bx = Ty * Nz - Tz * Ny; q.By = Tz * Nx - Tx * Nz; q.Bz = Tx * Ny - Ty * Nx;
  return q;
}

// Whole-disk decision.  Every disk sample lies on the plane through the sun
// normal to N within R_sun of the sun, so its direction is within
// atan(R_sun / |sun - p|) <= alpha = R_sun / |sun - p| of N (the larger
// alpha only widens the bound): its zenith angle within alpha of N's, its
// azimuth within alpha / min sin(theta) over [theta_c - alpha, theta_c +
// alpha].  The thresholds h[i] + t dphi it can meet then lie in [min h,
// max h + dphi] over the azimuth bins covered.  When the whole disk is below
// (above) that band by a margin covering FP64 rounding, every (no) sample
// counts, exactly as the per-sample loop would decide: returns ndisk - 1
// (0).  Otherwise returns -1 and the samples decide.
__device__ __forceinline__ int direct_decide(const DirectPair &q, const double *F,
                                             const double *h, int nphi, int ndisk) {
  const double TWO_PI = 2 * M_PI;
  const double dphi = TWO_PI / (nphi - 1);
  const double Nx = q.Nx, Ny = q.Ny, Nz = q.Nz;
  const double alpha = SUN_RADIUS_KM / q.nN;
  const double cz = (F[6] * Nx + F[7] * Ny) + F[8] * Nz;
  const double thc = acos(fmin(1.0, fmax(-1.0, cz)));
  const double smin = fmin(sin(thc - alpha), sin(thc + alpha));
  if (!(thc - alpha > 1e-3 && thc + alpha < M_PI - 1e-3 && smin > 0.0)) return -1;
  const double dph = alpha / smin * 1.0001 + 1e-9;
  if (!(dph < 0.25)) return -1;
  const double cx = (F[0] * Nx + F[1] * Ny) + F[2] * Nz;
  const double cy = (F[3] * Nx + F[4] * Ny) + F[5] * Nz;
  double phc = atan2(cy, cx);
  if (phc < 0) phc = phc + TWO_PI;
  const int nb = nphi - 1;  // bins of [0, 2 pi)
  const int i0 = (int)floor((phc - dph) / dphi) - 1, i1 = (int)floor((phc + dph) / dphi) + 1;
  double hlo = __builtin_huge_valf(), hhi = -__builtin_huge_valf();
  bool nan = false;
  // bin of step i (i0 <= i <= i1 + 1), or -1; i1 + 1: the azimuth 2 pi
  // itself maps to h[nphi - 1] when the range reaches the wrap
  auto bin = [&](int i) {
    int ii = i;
    if (i <= i1) {
      while (ii < 0) ii += nb;
      while (ii >= nb) ii -= nb;
    } else {
      ii = (i0 < 0 || i1 >= nb - 1) ? nb : -1;
    }
    return ii;
  };
  // the horizon samples are misses into a large array: the first
  // DIRECT_DECIDE_BATCH are loaded together, then min / max taken in order
  constexpr int DIRECT_DECIDE_BATCH = 8;
  for (int i = i0; i <= i1 + 1; i += DIRECT_DECIDE_BATCH) {
    double hv[DIRECT_DECIDE_BATCH];
#pragma unroll
    for (int u = 0; u < DIRECT_DECIDE_BATCH; ++u) {
      const int ii = i + u <= i1 + 1 ? bin(i + u) : -1;
      hv[u] = ii >= 0 ? h[ii] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < DIRECT_DECIDE_BATCH; ++u) {
      if (!(i + u <= i1 + 1 && bin(i + u) >= 0)) continue;
      nan = nan || hv[u] != hv[u];
      hlo = fmin(hlo, hv[u]);
      hhi = fmax(hhi, hv[u]);
    }
  }
  const double eps = 1e-9;
  if (!nan && thc + alpha + eps < hlo) return ndisk - 1;
  if (!nan && thc - alpha - eps >= hhi + dphi) return 0;
  return -1;
}

// Whether disk sample k of pair q is above the facet's horizon.
__device__ __forceinline__ bool direct_sample(const DirectPair &q, int k, const double *F,
                                              const double *h, int nphi,
                                              const double *__restrict__ disk, int ndisk) {
  const double TWO_PI = 2 * M_PI;
  const double dphi = TWO_PI / (nphi - 1);
  const double xk = disk[k], yk = disk[ndisk + k];
  const double dx = q.sx + SUN_RADIUS_KM * (xk * q.Tx + yk * q.Bx);
  const double dy = q.sy + SUN_RADIUS_KM * (xk * q.Ty + yk * q.By);
  const double dz = q.sz + SUN_RADIUS_KM * (xk * q.Tz + yk * q.Bz);
  double wx = dx - q.px, wy = dy - q.py, wz = dz - q.pz;
  const double nw = sqrt((wx * wx + wy * wy) + wz * wz);
  wx = wx / nw; wy = wy / nw; wz = wz / nw;
  const double d0 = (F[0] * wx + F[1] * wy) + F[2] * wz;
  const double d1 = (F[3] * wx + F[4] * wy) + F[5] * wz;
  const double d2 = (F[6] * wx + F[7] * wy) + F[8] * wz;
  const double theta = acos(d2);
  double phi = atan2(d1, d0);
  if (phi < 0) phi = phi + TWO_PI;
  int pi_ = (int)floor(phi / dphi);
  pi_ = pi_ < 0 ? 0 : (pi_ > nphi - 1 ? nphi - 1 : pi_);
  const double t = (phi - pi_ * dphi) / dphi;
  return theta < h[pi_] + t * dphi;
}

// E of pair q from its count of disk samples above the horizon.
__device__ __forceinline__ void direct_store(const DirectPair &q, int count, const double *F,
                                             int nphi, int64_t nf, double *__restrict__ E,
                                             const double *__restrict__ flux) {
  double E0 = (double)count / (nphi - 1);
  E0 = E0 * fmax(0.0, (q.Nx * F[6] + q.Ny * F[7]) + q.Nz * F[8]);
  const double ns = sqrt((q.sx * q.sx + q.sy * q.sy) + q.sz * q.sz);
  const double r = ONE_AU_KM / ns;
  E0 = E0 * (1372 * (r * r));
  // flux: the time loop's surface flux F_surf; its radiosity is
  // rad = -F_surf + direct (illum_cli.cpp:411-423), which equals
  // direct - F_surf exactly
  E[(int64_t)q.s * nf + q.f] = flux ? E0 - flux[q.f] : E0;
}

template <int L>
__global__ __launch_bounds__(256) void direct_kernel(
    int64_t nf, int S, int nphi, const float4 *__restrict__ pr4,
    const double *__restrict__ frame, const double *__restrict__ hz,
    const double *__restrict__ sun, const double *__restrict__ disk, int ndisk,
    double *__restrict__ E, int exact_only, const int32_t *__restrict__ body_of,
    int per_body, const int *__restrict__ step, const double *__restrict__ flux) {
  static_assert(L >= 1 && L <= 64 && (L & (L - 1)) == 0, "L divides the wavefront");
  const int64_t lane_id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t idx = lane_id / L;
  const int sub = (int)(lane_id % L);
  // a group of L lanes is all in range or all out of it, so the shuffles
  // below only read lanes of the caller's own (active) group
  if (idx >= nf * S) return;
  const DirectPair q = direct_setup(idx, nf, S, pr4, sun, body_of, per_body, step);
  const double *F = frame + 9 * q.f;
  const double *h = hz + q.f * nphi;
  const int dec = exact_only ? -1 : direct_decide(q, F, h, nphi, ndisk);
  int count = dec >= 0 && sub == 0 ? dec : 0;
  if (dec < 0)
    for (int k = sub; k < ndisk - 1; k += L) count += direct_sample(q, k, F, h, nphi, disk, ndisk);
  for (int o = L / 2; o > 0; o >>= 1) count += __shfl_xor(count, o);
  if (sub != 0) return;
  direct_store(q, count, F, nphi, nf, E, flux);
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int P>
__global__ __launch_bounds__(256) void direct_wave_kernel(
    int64_t nf, int S, int nphi, const float4 *__restrict__ pr4,
    const double *__restrict__ frame, const double *__restrict__ hz,
    const double *__restrict__ sun, const double *__restrict__ disk, int ndisk,
    double *__restrict__ E, int exact_only, const int32_t *__restrict__ body_of,
    int per_body, const int *__restrict__ step, const double *__restrict__ flux) {
  static_assert(P >= 1 && P <= 64 && (P & (P - 1)) == 0, "P divides the wavefront");
  const int64_t lane_id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t idx = lane_id / P;
  const bool owner = lane_id % P == 0;  // the lane that stores the pair's E
  const int lane = threadIdx.x & 63;
  const bool valid = idx < nf * S;
  // a lane past the last pair sets up pair 0 and takes no further part
  const DirectPair q = direct_setup(valid ? idx : 0, nf, S, pr4, sun, body_of, per_body, step);
  const double *F = frame + 9 * q.f;
  const double *h = hz + q.f * nphi;
  int count = 0;
  bool open = false;
  if (valid) {
    const int dec = exact_only ? -1 : direct_decide(q, F, h, nphi, ndisk);
    open = dec < 0;
    count = open ? 0 : dec;
  }
  // the open pairs, one after the other, each sampled by every lane
  uint64_t m = __builtin_amdgcn_ballot_w64(open && owner);
  while (m) {
    const int l = (int)__builtin_ctzll(m);
    m &= m - 1;
    DirectPair o;
    o.px = readlane_d(q.px, l); o.py = readlane_d(q.py, l); o.pz = readlane_d(q.pz, l);
    o.sx = readlane_d(q.sx, l); o.sy = readlane_d(q.sy, l); o.sz = readlane_d(q.sz, l);
    o.Tx = readlane_d(q.Tx, l); o.Ty = readlane_d(q.Ty, l); o.Tz = readlane_d(q.Tz, l);
    o.Bx = readlane_d(q.Bx, l); o.By = readlane_d(q.By, l); o.Bz = readlane_d(q.Bz, l);
    const int64_t of = __builtin_amdgcn_readlane((int)q.f, l);
    const double *oF = frame + 9 * of;
    const double *oh = hz + of * nphi;
    int c = 0;
    for (int k = lane; k < ndisk - 1; k += 64) c += direct_sample(o, k, oF, oh, nphi, disk, ndisk);
    for (int w = 32; w > 0; w >>= 1) c += __shfl_xor(c, w);
    if (lane == l) count = c;
  }
  if (valid && owner) direct_store(q, count, F, nphi, nf, E, flux);
}

// Two-pass shape.  Pass 1 (direct_split_kernel): one lane per pair does the
// set-up and the whole-disk decision and stores the decided pairs; the open
// ones go to the mesh's queue (one atomic per wavefront), or, past its
// capacity, are sampled by their own lane.  Pass 2 (direct_queue_kernel): a
// fixed grid in which each wavefront takes queued pairs in turn and samples
// one with all 64 lanes.  The open pairs' samples so run 64 to a wavefront
// wherever they are, where one lane group per pair (direct_kernel) leaves the
// rest of the wavefront idle and direct_wave_kernel serialises a wavefront's
// open pairs.  The last workgroup of pass 2 to finish zeroes the queue's
// count and its own counter for the next launch, so no clear is queued (the
// launches may be captured into a graph).  Counts are integers: E is the
// same as with the other shapes.
constexpr int DIRECT_QUEUE_BLOCKS = 1024;

__global__ __launch_bounds__(256) void direct_split_kernel(
    int64_t nf, int S, int nphi, const float4 *__restrict__ pr4,
    const double *__restrict__ frame, const double *__restrict__ hz,
    const double *__restrict__ sun, const double *__restrict__ disk, int ndisk,
    double *__restrict__ E, int exact_only, const int32_t *__restrict__ body_of,
    int per_body, const int *__restrict__ step, const double *__restrict__ flux,
    int32_t *__restrict__ dq, int64_t dq_cap) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool valid = idx < nf * S;
  const DirectPair q = direct_setup(valid ? idx : 0, nf, S, pr4, sun, body_of, per_body, step);
  const double *F = frame + 9 * q.f;
  const double *h = hz + q.f * nphi;
  int dec = 0;
  if (valid) dec = exact_only ? -1 : direct_decide(q, F, h, nphi, ndisk);
  const bool open = valid && dec < 0;
  const uint64_t m = __builtin_amdgcn_ballot_w64(open);
  int base = 0;
  if (m) {
    if (lane == (int)__builtin_ctzll(m)) base = atomicAdd(&dq[0], (int)__popcll(m));
    base = __builtin_amdgcn_readlane(base, (int)__builtin_ctzll(m));
  }
  if (!valid) return;
  if (open) {
    const int64_t slot = (int64_t)base + __popcll(m & ((1ull << lane) - 1));
    if (slot < dq_cap) {
      dq[2 + slot] = (int32_t)idx;
      return;
    }
    int c = 0;  // past the queue's capacity: this lane samples its pair
    for (int k = 0; k < ndisk - 1; ++k) c += direct_sample(q, k, F, h, nphi, disk, ndisk);
    dec = c;
  }
  direct_store(q, dec, F, nphi, nf, E, flux);
}

__global__ __launch_bounds__(256) void direct_queue_kernel(
    int64_t nf, int S, int nphi, const float4 *__restrict__ pr4,
    const double *__restrict__ frame, const double *__restrict__ hz,
    const double *__restrict__ sun, const double *__restrict__ disk, int ndisk,
    double *__restrict__ E, const int32_t *__restrict__ body_of, int per_body,
    const int *__restrict__ step, const double *__restrict__ flux, int32_t *__restrict__ dq,
    int64_t dq_cap) {
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t n = std::min<int64_t>(__builtin_amdgcn_readfirstlane(dq[0]), dq_cap);
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t e = (int64_t)blockIdx.x * 4 + wid; e < n; e += nw) {
    const int64_t idx = __builtin_amdgcn_readfirstlane(dq[2 + e]);
    const DirectPair q = direct_setup(idx, nf, S, pr4, sun, body_of, per_body, step);
    const double *F = frame + 9 * q.f;
    const double *h = hz + q.f * nphi;
    int c = 0;
    for (int k = lane; k < ndisk - 1; k += 64) c += direct_sample(q, k, F, h, nphi, disk, ndisk);
    for (int w = 32; w > 0; w >>= 1) c += __shfl_xor(c, w);
    if (lane == 0) direct_store(q, c, F, nphi, nf, E, flux);
  }
  // every wavefront has read the count before its workgroup arrives here
  __syncthreads();
  if (threadIdx.x == 0 && atomicAdd(&dq[1], 1) == (int)gridDim.x - 1) {
    atomicExch(&dq[0], 0);
    atomicExch(&dq[1], 0);
  }
}

// One direct insolation launch over nf x S pairs: L > 0 lanes per pair
// (direct_kernel<L>), or direct_wave_kernel<P> for L = 0 (P = 1) and
// L = -4, -8, -16 (P = -L), or the two-pass shape for L = DIRECT_SPLIT
// (below 2^31 pairs: the queue holds 32-bit pair indices).  The shape does
// not change E (the lanes only share integer counts), so it is picked by the
// autotuner per pair-count scale (ILLUM_DIRECT_LANES=-16|-8|-4|0|1|2|4|8|16|64
// overrides).  The trials
// cover the first DIRECT_TUNE_PAIRS pairs only (a grid over a prefix of the
// pairs writes exactly their E, which the launch after the trials writes
// again): a batch of 64 bodies x 32 suns has 1.7e8 pairs, and timing nine
// shapes twice over all of them cost several times the launch itself.
constexpr int64_t DIRECT_TUNE_PAIRS = (int64_t)1 << 21;

static void direct_launch(illum_mesh *mesh, int nphi, const double *hz, const double *sun, int S,
                          const double *disk, int ndisk, double *E, bool per_body,
                          const int *step, const double *flux, hipStream_t s) {
  const int64_t nf = mesh->g.nf;
  const int64_t pairs = nf * S;
  auto go_n = [&](int L, int64_t npairs) {
    if (L == DIRECT_SPLIT) {
      direct_split_kernel<<<(unsigned)((npairs + 255) / 256), 256, 0, s>>>(
          nf, S, nphi, (const float4 *)mesh->d_pr4, mesh->d_frame, hz, sun, disk, ndisk, E,
          direct_exact_only() ? 1 : 0, mesh->d_body, per_body ? 1 : 0, step, flux, mesh->d_dq,
          mesh->dq_cap);
      ILLUM_LAUNCH_CHECK();
      direct_queue_kernel<<<DIRECT_QUEUE_BLOCKS, 256, 0, s>>>(
          nf, S, nphi, (const float4 *)mesh->d_pr4, mesh->d_frame, hz, sun, disk, ndisk, E,
          mesh->d_body, per_body ? 1 : 0, step, flux, mesh->d_dq, mesh->dq_cap);
      ILLUM_LAUNCH_CHECK();
      return;
    }
    const int64_t lanes = npairs * (L > 0 ? L : L == 0 ? 1 : -L);
    auto kern = L == 0 ? direct_wave_kernel<1> : L == -4 ? direct_wave_kernel<4>
              : L == -8 ? direct_wave_kernel<8> : L == -16 ? direct_wave_kernel<16>
              : L == 16 ? direct_kernel<16>
              : L == 8 ? direct_kernel<8> : L == 4 ? direct_kernel<4>
              : L == 2 ? direct_kernel<2> : direct_kernel<1>;
    kern<<<(unsigned)((lanes + 255) / 256), 256, 0, s>>>(
        nf, S, nphi, (const float4 *)mesh->d_pr4, mesh->d_frame, hz, sun, disk, ndisk, E,
        direct_exact_only() ? 1 : 0, mesh->d_body, per_body ? 1 : 0, step, flux);
    ILLUM_LAUNCH_CHECK();
  };
  const bool split_ok = pairs <= INT32_MAX && mesh->d_dq;
  int L = direct_lanes(pairs);
  if (L == DIRECT_SPLIT && !split_ok) L = 16;
  if (!getenv("ILLUM_DIRECT_LANES")) {
    int b = 0;
    while (b < 62 && ((int64_t)1 << (b + 1)) <= pairs) ++b;
    const int64_t sample = std::min(pairs, DIRECT_TUNE_PAIRS);
    std::vector<int> cands = {-16, -8, -4, 0, 1, 2, 4, 8, 16};
    if (split_ok) cands.push_back(DIRECT_SPLIT);
    L = autotune(mesh->ctx, "direct_lanes[pairs~2^" + std::to_string(b) + "]", cands, L,
                 [&](int c) { go_n(c, sample); });
    if (L == DIRECT_SPLIT && !split_ok) L = 16;
  }
  go_n(L, pairs);
}

void direct_radiosity(illum_mesh *mesh, const double *hz, int nphi, const double *sun, int S,
                      const double *disk, int ndisk, double *E, int flags) {
  illum_ctx *ctx = mesh->ctx;
  hipStream_t s = stream_of(ctx);
  const int64_t nf = mesh->g.nf;
  if (nphi < 2 || S < 1 || ndisk < 1) throw Error(ILLUM_ERR_ARG, "bad direct radiosity sizes");
  if (!hz || !sun || !disk || !E) throw Error(ILLUM_ERR_ARG, "null pointer");
  const bool in_dev = flags & ILLUM_IN_ON_DEVICE, out_dev = flags & ILLUM_OUT_ON_DEVICE;
  const bool per_body = flags & ILLUM_SUN_PER_BODY;
  const int64_t nsun = per_body ? (int64_t)S * mesh->nbody : S;
  DBuf<double> d_h, d_sun, d_disk, d_E;
  const double *ph = hz, *psun = sun, *pdisk = disk;
  if (!in_dev) {
    d_h.reset(nf * nphi); d_sun.reset(3 * nsun); d_disk.reset(2 * ndisk);
    ILLUM_HIP(hipMemcpyAsync(d_h.p, hz, 8 * nf * nphi, hipMemcpyHostToDevice, s));
    ILLUM_HIP(hipMemcpyAsync(d_sun.p, sun, 8 * 3 * nsun, hipMemcpyHostToDevice, s));
    ILLUM_HIP(hipMemcpyAsync(d_disk.p, disk, 8 * 2 * ndisk, hipMemcpyHostToDevice, s));
    ph = d_h.p; psun = d_sun.p; pdisk = d_disk.p;
  }
  double *pE = E;
  if (!out_dev) { d_E.reset(nf * S); pE = d_E.p; }
  {
    StageTimer st(ctx, "direct");
    direct_launch(mesh, nphi, ph, psun, S, pdisk, ndisk, pE, per_body, nullptr, nullptr, s);
  }
  if (!out_dev) ILLUM_HIP(hipMemcpyAsync(E, pE, 8 * nf * S, hipMemcpyDeviceToHost, s));
  if (!(in_dev && out_dev && (flags & ILLUM_ASYNC))) sync(ctx);
}

// Queues direct insolation for one step of a time loop on stream s, all
// arrays on the device: sun column (*step) S .. + S - 1, E = direct - flux.
// No host synchronisation (graph-capturable; the lanes per pair must have
// been tuned by an earlier call outside the capture).
void direct_enqueue(illum_mesh *mesh, const double *hz, int nphi, const double *sun, int S,
                    const double *disk, int ndisk, double *E, const int *step, const double *flux,
                    hipStream_t s) {
  direct_launch(mesh, nphi, hz, sun, S, disk, ndisk, E, false, step, flux, s);
}

}  // namespace illum
