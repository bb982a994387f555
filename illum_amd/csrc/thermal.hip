// Batched per-facet Crank-Nicolson subsurface conduction
// (thermal_model::step, thermal.hpp:65-105 -> conductionQ, conductionQ.f:35-98,
// tridiagonal solve per tridag.for:17-27, i.e. the Thomas algorithm; see
// Golub & Van Loan, Matrix Computations, section 4.3).
//
// Layout.  Facets are independent.  They are permuted into slots so that each
// 64-slot wavefront holds facets of a single layer grid (each grid's list is
// padded to a multiple of 64 with empty slots); per-layer state is slot-major,
// T[l * nslot + slot] (l = 0 surface, 1..nz layers), so state loads and
// stores are coalesced and a wavefront's grid coefficients are uniform.
//
// Engines.  Grids with at most NZ layers (NZ in 32/64/96/128) run on
// register kernels: a lane keeps its whole temperature column in registers
// for every step of the launch.  Deeper grids (<= 1000 layers,
// conductionQ.f:38 NMAX) run on cn_stream_kernel, whose column streams
// through HBM each step (coalesced across the wavefront).
// ILLUM_CN_ENGINE=stream forces the streaming kernel for comparison.
//
// Coefficients: the time-invariant tridiagonal entries of each grid are
// tabulated on the host in conductionQ's operation order for the current dt; a
// facet's thermal-inertia scale s multiplies every conductivity by s^2.
//
// FP64 mode (cn_fp64_kernel): the reference step verbatim -- right-hand side,
// Thomas sweep with one division per row, surface temperature and flux, all
// in FP64.
// Mixed mode (cn_mixed_kernel): the increment d = T^{n+1} - T^n is solved in
// FP32 against an FP64 residual (see the kernel), T^{n+1} = T^n + d in FP64.
// |d| is the per-step temperature change, so the FP32 error scales with the
// change, not with T.  The radiative surface linearisation and the surface
// flux stay FP64.  The streaming kernel's mixed mode forms rho = r - A T^n
// directly and runs a top-down FP32 Thomas sweep.
#include "hip_util.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <vector>

namespace illum {

constexpr double SIGMA_SB = 5.6704e-8;  // conductionQ.f:39

// Per-grid header (6 doubles): k1 = k(1)/dz, k(1), z(1), dt/rhoc(nz),
// z(nz) - z(nz-1), nz.  Per-layer table: interleaved alpha, gamma.
struct GridHdr {
  double k1, kfirst, z1, geo, dzlast, nz;
};

// Linearised radiative boundary at the old surface temperature
// (conductionQ.f:76-82): returns ann, annp1, bn.
__device__ __forceinline__ void surface_bc(double Tsurf, double Qn, double Qnp1, double k1,
                                           double emiss, double *ann, double *annp1,
                                           double *bn) {
  const double Tr = Tsurf;
  const double Tr3 = (Tr * Tr) * Tr, Tr4 = Tr3 * Tr;  // Tr**3, Tr**4 left to right
  const double arad = ((-3.0 * emiss) * SIGMA_SB) * Tr4;
  const double brad = ((2.0 * emiss) * SIGMA_SB) * Tr3;
  const double den = k1 + brad;
  *ann = (Qn - arad) / den;
  *annp1 = (Qnp1 - arad) / den;
  *bn = (k1 - brad) / den;
}

// Waves per SIMD of a register kernel: columns of up to 32 layers fit two
// waves' register files; deeper columns take the whole file (512 VGPR+AGPR).
constexpr int cn_waves(int nz) { return nz <= 32 ? 2 : 1; }

// ---------------------------------------------------------------------------
// FP64 register engine (the reference step verbatim): the temperature column
// of a lane stays in registers for all steps of the launch.  The Thomas
// multipliers stay in registers too when GMREG (columns of up to 96 layers);
// otherwise they go to a wavefront-private scratch block, written in the
// forward sweep and read back in the backward sweep.
template <int NZ, bool GMREG>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(cn_waves(NZ), cn_waves(NZ)))) void cn_fp64_kernel(
    const int32_t *__restrict__ fidx, const int32_t *__restrict__ wgrid, int64_t block0,
    int64_t nf, int64_t nslot, const double *__restrict__ s2f, const double *__restrict__ hdr,
    const double *__restrict__ coef, const int64_t *__restrict__ goff, double emiss, double Fgeo,
    int nsteps, int64_t qs, const double *__restrict__ Q, double *__restrict__ Qprev,
    double *__restrict__ T, double *__restrict__ Fs, double *__restrict__ G,
    double *__restrict__ Tsurf_series, int *__restrict__ status, int rows) {
  __shared__ double s_al[NZ], s_ga[NZ], s_geo[NZ];
  const int64_t blk = block0 + blockIdx.x;
  const int64_t slot = blk * 64 + threadIdx.x;
  const int f = fidx[slot];
  const bool act = f >= 0;
  const int g = wgrid[blk];
  const GridHdr H = *(const GridHdr *)(hdr + 6 * g);
  const int nz = (int)H.nz;  // uniform across the wavefront, <= NZ
  const double *al = coef + 2 * goff[g];
  const double geo_term = (H.geo * Fgeo) / H.dzlast;
  // rows past nz: zero coefficients (and the geothermal term on row nz only)
  for (int j = threadIdx.x; j < NZ; j += 64) {
    s_al[j] = j < nz ? al[2 * j] : 0.0;
    s_ga[j] = j < nz ? al[2 * j + 1] : 0.0;
    s_geo[j] = j == nz - 1 ? geo_term : 0.0;
  }
  __syncthreads();
  const double s2 = (act && s2f) ? s2f[f] : 1.0;
  const double k1 = s2 * H.k1, kfirst = s2 * H.kfirst;
  // Wavefront-private scratch block, element [row][lane]: the per-layer offset
  // is a compile-time immediate on one base address.
  double *Gw = GMREG ? nullptr : G + blk * (int64_t)rows * 64 + threadIdx.x;
  double Gm[GMREG ? NZ + 2 : 1];

  double Tl[NZ + 2];  // Tl[0] surface, Tl[1..nz] layers, one spare slot
#pragma unroll
  for (int l = 0; l < NZ + 2; ++l) Tl[l] = (act && l <= nz) ? T[(int64_t)l * nslot + slot] : 0.0;
  double Qn = act ? Qprev[f] : 0.0;
  double Fsurf = act ? Fs[f] : 0.0;
  bool zero_pivot = false;

  for (int step = 0; step < nsteps; ++step) {
    const double Qnp1 = act ? Q[(int64_t)step * qs + f] : 0.0;
    double ann, annp1, bn;
    surface_bc(Tl[0], Qn, Qnp1, k1, emiss, &ann, &annp1, &bn);
    const double a1 = s2 * s_al[0], g1 = s2 * s_ga[0];
    const double b1 = ((1.0 + a1) + g1) - g1 * bn;
    if (b1 == 0.0) zero_pivot = true;
    double bet = 0.0, uprev = 0.0, cprev = 0.0, Told = 0.0;
    // With the multipliers in registers (NZ <= 96) the sweeps run over all NZ
    // rows without branches.  Rows past nz have zero coefficients and T = 0,
    // so they stay 0 and, the bottom row's alpha being 0, never reach row nz.
    // The interior formula with alpha = 0 plus the geothermal term is the
    // bottom row's r and b term for term, so every real row is computed
    // exactly as the reference's tridag step computes it.
#pragma unroll
    for (int j = 1; j <= NZ; ++j) {
      if (GMREG || j <= nz) {
        const double Tc = Tl[j];
        const double Tn = (GMREG || j < nz) ? Tl[j + 1] : 0.0;
        const double aj = j == 1 ? a1 : s2 * s_al[j - 1];
        const double gj = j == 1 ? g1 : s2 * s_ga[j - 1];
        double r, bj;
        if (j == 1) {
          r = (g1 * (annp1 + ann) + (((1.0 - a1) - g1) + g1 * bn) * Tc) + a1 * Tn;
          bj = b1;
        } else if (GMREG) {
          r = ((gj * Told + ((1.0 - aj) - gj) * Tc) + aj * Tn) + s_geo[j - 1];
          bj = (1.0 + aj) + gj;
        } else if (j < nz) {
          r = (gj * Told + ((1.0 - aj) - gj) * Tc) + aj * Tn;
          bj = (1.0 + aj) + gj;
        } else {
          r = (gj * Told + (1.0 - gj) * Tc) + geo_term;
          bj = 1.0 + gj;
        }
        double u;
        if (j == 1) {
          bet = bj;
          u = r / bet;
        } else {
          const double a_j = -gj;
          const double gm = cprev / bet;
          if (GMREG) Gm[j] = gm;
          else Gw[j * 64] = gm;
          bet = bj - a_j * gm;
          u = (r - a_j * uprev) / bet;
        }
        Tl[j] = u;  // old T_j is kept in Told for the next row
        uprev = u;
        cprev = -aj;
        Told = Tc;
      }
      // Keep the scheduler from hoisting loads across groups of 8 layers,
      // which bounds register pressure (the column itself needs 2 VGPRs per layer).
      if ((j & 7) == 0) __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int j = NZ - 1; j >= 1; --j) {
      if (GMREG || j < nz) Tl[j] = Tl[j] - (GMREG ? Gm[j + 1] : Gw[(j + 1) * 64]) * Tl[j + 1];
      if ((j & 7) == 0) __builtin_amdgcn_sched_barrier(0);
    }
    const double T1 = Tl[1];
    const double Tsurf = 0.5 * ((annp1 + bn * T1) + T1);
    Fsurf = ((-kfirst) * (T1 - Tsurf)) / H.z1;
    Tl[0] = Tsurf;
    if (Tsurf_series && act) Tsurf_series[(int64_t)step * nf + f] = Tsurf;
    Qn = Qnp1;
  }
  if (act) {
#pragma unroll
    for (int l = 0; l < NZ + 2; ++l)
      if (l <= nz) T[(int64_t)l * nslot + slot] = Tl[l];
    Fs[f] = Fsurf;
    Qprev[f] = Qn;  // the last step's Q, the next call's Q^n
    if (zero_pivot) atomicOr(status, 1);
  }
}

// ---------------------------------------------------------------------------
// Mixed-precision register engine.  A step A T^{n+1} = r(T^n) is solved in
// increment form, A d = rho with rho = r - A T^n and T^{n+1} = T^n + d.
// For rows j >= 2 rho reduces to the explicit diffusion term
//   rho_j = 2 (alpha_j (T_{j+1} - T_j) - gamma_j (T_j - T_{j-1}))  (+ geothermal
// term on the bottom row), formed from FP64 differences of T so it carries the
// per-step change at full relative precision; row 1 keeps the linearised
// radiative terms in FP64.  The FP32 solve factors A from the bottom row up:
//   beta_nz = b_nz,  beta_j = b_j - alpha_j gamma_{j+1} / beta_{j+1},
//   z_j = (rho_j + alpha_j z_{j+1}) / beta_j,
//   d_1 = z_1,  d_j = z_j + (gamma_j / beta_j) d_{j-1},
// so only row 1 (whose diagonal follows the surface temperature) differs
// between steps.  Everything of rows >= 2 is therefore computed once per
// launch, per lane (the facet's conductivity factor s2 scales alpha and
// gamma): with P'_j = s2 / beta_j and the grid's unscaled alpha, gamma,
//   Q_j = alpha_j P'_j,  R_j = gamma_j P'_j,  e_j = T_j - T_{j-1},
//   z_j = Q_j (z_{j+1} + e_{j+1}) - R_j e_j,
//   d_j = R_j d_{j-1} + z_j.
// A step then costs per layer one FP64 difference and one FP64 update, one
// add, one multiply and two fused multiply-adds in FP32, with no coefficient
// loads.  The column T (FP64), z, Q and R (FP32) stay in registers.  rho is
// carried halved (the factor 2 is applied in the FP64 update).  GEO adds the
// geothermal term on the bottom row (the thermal model's flux is 0,
// thermal.hpp:79, so the usual instantiation leaves it out).
template <int NZ, bool GEO>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(cn_waves(NZ), cn_waves(NZ)))) void cn_mixed_kernel(
    const int32_t *__restrict__ fidx, const int32_t *__restrict__ wgrid, int64_t block0,
    int64_t nf, int64_t nslot, const double *__restrict__ s2f, const double *__restrict__ hdr,
    const double *__restrict__ coef, const float4 *__restrict__ coefp,
    const int64_t *__restrict__ goff, const int64_t *__restrict__ goffp, double emiss,
    double Fgeo, int nsteps, int64_t qs, const double *__restrict__ Q, double *__restrict__ Qprev,
    double *__restrict__ T, double *__restrict__ Fs, double *__restrict__ Tsurf_series) {
  const int64_t blk = block0 + blockIdx.x;
  const int64_t slot = blk * 64 + threadIdx.x;
  const int f = fidx[slot];
  const bool act = f >= 0;
  const int g = wgrid[blk];
  const GridHdr H = *(const GridHdr *)(hdr + 6 * g);
  const int nz = (int)H.nz;  // uniform across the wavefront, <= NZ
  const double *al = coef + 2 * goff[g];
  // per layer j (entry j - 1): alpha, gamma, alpha + gamma, 1 on the bottom row;
  // zeros past the grid's nz up to NZ (the bucket's depth)
  const float4 *cq = coefp + goffp[g];
  const double s2 = (act && s2f) ? s2f[f] : 1.0;
  const float s2s = (float)s2;
  const double k1 = s2 * H.k1, kfirst = s2 * H.kfirst;
  const float geo_half = (float)(0.5 * ((H.geo * Fgeo) / H.dzlast));
  const double a1 = s2 * al[0], g1 = s2 * al[1];
  const float a1s = (float)a1;

  double Tl[NZ + 1];  // Tl[0] surface, Tl[1..nz] layers
#pragma unroll
  for (int l = 0; l <= NZ; ++l) Tl[l] = (act && l <= nz) ? T[(int64_t)l * nslot + slot] : 0.0;
  double Qn = act ? Qprev[f] : 0.0;
  double Fsurf = act ? Fs[f] : 0.0;

  // Time-invariant factors of rows nz..2 (zero for rows past nz, so those
  // rows give z_j = d_j = 0 and T_j stays 0; the bottom row's alpha is 0, so
  // the row below it does not reach it).
  float Qc[NZ + 1], Rc[NZ + 1];
  float Gb = 0.0f;  // GEO: the bottom row's geothermal term in z's scaling
  {
    float Mn = 0.0f;  // gamma_{j+1} P'_{j+1}
#pragma unroll
    for (int j = NZ; j >= 2; --j) {
      if (j <= nz) {
        const float4 c = cq[j - 1];
        const float bj = __builtin_fmaf(s2s, c.z, 1.0f);
        const float beta = __builtin_fmaf(-s2s * c.x, Mn, bj);
        const float Pp = s2s * __builtin_amdgcn_rcpf(beta);
        Qc[j] = c.x * Pp;
        Rc[j] = c.y * Pp;
        if (GEO && c.w != 0.0f) Gb = (geo_half / s2s) * Pp;
        Mn = Rc[j];
      } else {
        Qc[j] = 0.0f;
        Rc[j] = 0.0f;
      }
    }
  }
  const float M2 = Rc[2];  // gamma_2 P'_2: enters row 1's pivot every step

  for (int step = 0; step < nsteps; ++step) {
    const double Qnp1 = act ? Q[(int64_t)step * qs + f] : 0.0;
    double ann, annp1, bn;
    surface_bc(Tl[0], Qn, Qnp1, k1, emiss, &ann, &annp1, &bn);
    float z[NZ + 1];
    float znext = 0.0f, enext = 0.0f;  // z and T difference of row j+1
#pragma unroll
    for (int j = NZ; j >= 2; --j) {
      const float ej = (float)(Tl[j] - Tl[j - 1]);
      float zj = __builtin_fmaf(Qc[j], znext + enext, -Rc[j] * ej);
      if (GEO && j == nz) zj += Gb;
      z[j] = zj;
      znext = zj;
      enext = ej;
      if ((j & 7) == 0) __builtin_amdgcn_sched_barrier(0);
    }
    {
      const double b1 = ((1.0 + a1) + g1) - g1 * bn;
      const double surf = (0.5 * g1) * ((ann + annp1) + (2.0 * (bn - 1.0)) * Tl[1]);
      const float rho1 = __builtin_fmaf(a1s, enext, (float)surf);
      z[1] = __builtin_fmaf(a1s, znext, rho1) *
             __builtin_amdgcn_rcpf(__builtin_fmaf(-a1s, M2, (float)b1));
    }
    float d = z[1];
    Tl[1] = __builtin_fma(2.0, (double)d, Tl[1]);
#pragma unroll
    for (int j = 2; j <= NZ; ++j) {
      d = __builtin_fmaf(Rc[j], d, z[j]);
      Tl[j] = __builtin_fma(2.0, (double)d, Tl[j]);
      if ((j & 7) == 0) __builtin_amdgcn_sched_barrier(0);
    }
    const double T1 = Tl[1];
    const double Tsurf = 0.5 * ((annp1 + bn * T1) + T1);
    Fsurf = ((-kfirst) * (T1 - Tsurf)) / H.z1;
    Tl[0] = Tsurf;
    if (Tsurf_series && act) Tsurf_series[(int64_t)step * nf + f] = Tsurf;
    Qn = Qnp1;
  }
  if (act) {
#pragma unroll
    for (int l = 0; l <= NZ; ++l)
      if (l <= nz) T[(int64_t)l * nslot + slot] = Tl[l];
    Fs[f] = Fsurf;
    Qprev[f] = Qn;
  }
}

// The same step with the per-lane pivots P'_j kept and the grid's alpha,
// gamma reloaded every step by scalar loads (one register array fewer than
// cn_mixed_kernel): z_j = (alpha_j P'_j) z_{j+1} + (alpha_j e_{j+1} -
// gamma_j e_j) P'_j, d_j = z_j + gamma_j P'_j d_{j-1}.  Kept for comparison
// (ILLUM_CN_MIXED_FORM=1).
template <int NZ>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(cn_waves(NZ), cn_waves(NZ)))) void cn_mixed_p_kernel(
    const int32_t *__restrict__ fidx, const int32_t *__restrict__ wgrid, int64_t block0,
    int64_t nf, int64_t nslot, const double *__restrict__ s2f, const double *__restrict__ hdr,
    const double *__restrict__ coef, const float4 *__restrict__ coefp,
    const int64_t *__restrict__ goff, const int64_t *__restrict__ goffp, double emiss,
    double Fgeo, int nsteps, int64_t qs, const double *__restrict__ Q, double *__restrict__ Qprev,
    double *__restrict__ T, double *__restrict__ Fs, double *__restrict__ Tsurf_series) {
  const int64_t blk = block0 + blockIdx.x;
  const int64_t slot = blk * 64 + threadIdx.x;
  const int f = fidx[slot];
  const bool act = f >= 0;
  const int g = wgrid[blk];
  const GridHdr H = *(const GridHdr *)(hdr + 6 * g);
  const int nz = (int)H.nz;  // uniform across the wavefront, <= NZ
  const double *al = coef + 2 * goff[g];
  // per layer j (entry j - 1): alpha, gamma, alpha + gamma, 1 on the bottom row;
  // zeros past the grid's nz up to NZ (the bucket's depth)
  const float4 *cq = coefp + goffp[g];
  const double s2 = (act && s2f) ? s2f[f] : 1.0;
  const float s2s = (float)s2;
  const double k1 = s2 * H.k1, kfirst = s2 * H.kfirst;
  const float geo_half = (float)(0.5 * ((H.geo * Fgeo) / H.dzlast));
  const double a1 = s2 * al[0], g1 = s2 * al[1];
  const float a1s = (float)a1;

  double Tl[NZ + 1];  // Tl[0] surface, Tl[1..nz] layers
#pragma unroll
  for (int l = 0; l <= NZ; ++l) Tl[l] = (act && l <= nz) ? T[(int64_t)l * nslot + slot] : 0.0;
  double Qn = act ? Qprev[f] : 0.0;
  double Fsurf = act ? Fs[f] : 0.0;

  // Time-invariant pivots of rows nz..2, pre-scaled: P'_j = s2 / beta_j.
  float Pp[NZ + 1];
  float M2 = 0.0f;  // gamma_2 P'_2: enters row 1's pivot every step
  {
    float Mn = 0.0f;  // gamma_{j+1} P'_{j+1}
#pragma unroll
    for (int j = NZ; j >= 2; --j) {
      if (j <= nz) {
        const float4 c = cq[j - 1];
        const float bj = __builtin_fmaf(s2s, c.z, 1.0f);
        const float beta = __builtin_fmaf(-s2s * c.x, Mn, bj);
        Pp[j] = s2s * __builtin_amdgcn_rcpf(beta);
        Mn = c.y * Pp[j];
      } else {
        Pp[j] = 0.0f;  // rows past nz: z_j = d_j = 0, T_j stays 0
      }
    }
    M2 = Mn;
  }
  const float geo_z = geo_half / s2s;  // bottom-row geothermal term in z's scaling

  for (int step = 0; step < nsteps; ++step) {
    const double Qnp1 = act ? Q[(int64_t)step * qs + f] : 0.0;
    double ann, annp1, bn;
    surface_bc(Tl[0], Qn, Qnp1, k1, emiss, &ann, &annp1, &bn);
    // The sweeps run over all NZ rows without branches: rows past nz have
    // zero coefficients and pivots, so they contribute exact zeros, and the
    // bottom row's alpha is 0, so the row below it does not reach it.  The
    // coefficients are reloaded every step (scalar loads) rather than held in
    // scalar registers across the time loop.
    __asm__ volatile("" ::: "memory");
    float z[NZ + 1];
    float znext = 0.0f, enext = 0.0f;  // z and T difference of row j+1
#pragma unroll
    for (int j = NZ; j >= 2; --j) {
      const float4 c = cq[j - 1];
      const float ej = (float)(Tl[j] - Tl[j - 1]);
      float w = __builtin_fmaf(c.x, enext, -c.y * ej);
      w = __builtin_fmaf(geo_z, c.w, w);  // geothermal term on the bottom row
      // one fused multiply-add on the z recurrence; the rest overlaps
      z[j] = __builtin_fmaf(c.x * Pp[j], znext, w * Pp[j]);
      znext = z[j];
      enext = ej;
      if ((j & 7) == 0) __builtin_amdgcn_sched_barrier(0);
    }
    {
      const double b1 = ((1.0 + a1) + g1) - g1 * bn;
      const double surf = (0.5 * g1) * ((ann + annp1) + (2.0 * (bn - 1.0)) * Tl[1]);
      const float rho1 = __builtin_fmaf(a1s, enext, (float)surf);
      z[1] = __builtin_fmaf(a1s, znext, rho1) *
             __builtin_amdgcn_rcpf(__builtin_fmaf(-a1s, M2, (float)b1));
    }
    float d = z[1];
    Tl[1] = __builtin_fma(2.0, (double)d, Tl[1]);
#pragma unroll
    for (int j = 2; j <= NZ; ++j) {
      d = __builtin_fmaf(cq[j - 1].y * Pp[j], d, z[j]);
      Tl[j] = __builtin_fma(2.0, (double)d, Tl[j]);
      if ((j & 7) == 0) __builtin_amdgcn_sched_barrier(0);
    }
    const double T1 = Tl[1];
    const double Tsurf = 0.5 * ((annp1 + bn * T1) + T1);
    Fsurf = ((-kfirst) * (T1 - Tsurf)) / H.z1;
    Tl[0] = Tsurf;
    if (Tsurf_series && act) Tsurf_series[(int64_t)step * nf + f] = Tsurf;
    Qn = Qnp1;
  }
  if (act) {
#pragma unroll
    for (int l = 0; l <= NZ; ++l)
      if (l <= nz) T[(int64_t)l * nslot + slot] = Tl[l];
    Fs[f] = Fsurf;
    Qprev[f] = Qn;
  }
}

// ---------------------------------------------------------------------------
// Streaming engine (any depth): the column lives in HBM; scratch rows hold
// the Thomas multipliers (FP64) or the FP32 multipliers and forward values.
template <bool MIXED>
__global__ __launch_bounds__(64) void cn_stream_kernel(
    const int32_t *__restrict__ fidx, const int32_t *__restrict__ wgrid, int64_t block0,
    int64_t nf, int64_t nslot, const double *__restrict__ s2f, const double *__restrict__ hdr,
    const double *__restrict__ coef, const int64_t *__restrict__ goff, double emiss, double Fgeo,
    int nsteps, int64_t qs, const double *__restrict__ Q, double *__restrict__ Qprev,
    double *__restrict__ T, double *__restrict__ Fs, double *__restrict__ G,
    float *__restrict__ G32, double *__restrict__ Tsurf_series, int *__restrict__ status,
    int rows) {
  const int64_t blk = block0 + blockIdx.x;
  const int64_t slot = blk * 64 + threadIdx.x;
  const int f = fidx[slot];
  if (f < 0) return;
  // FP64 multipliers: this wavefront's own block of the scratch, element
  // [row][lane] (the layout of cn_fp64_kernel's scratch, so concurrently
  // running depth buckets never share scratch words)
  double *Gw = G ? G + blk * (int64_t)rows * 64 + threadIdx.x : nullptr;
  const int g = wgrid[blk];
  const GridHdr H = *(const GridHdr *)(hdr + 6 * g);
  const int nz = (int)H.nz;
  const double *al = coef + 2 * goff[g];
  const double s2 = s2f ? s2f[f] : 1.0;
  const double k1 = s2 * H.k1, kfirst = s2 * H.kfirst;
  const double geo_term = (H.geo * Fgeo) / H.dzlast;
  auto Tat = [&](int l) -> double & { return T[(int64_t)l * nslot + slot]; };
  double Qn = Qprev[f];
  double Fsurf = Fs[f];
  for (int step = 0; step < nsteps; ++step) {
    const double Qnp1 = Q[(int64_t)step * qs + f];
    double ann, annp1, bn;
    surface_bc(Tat(0), Qn, Qnp1, k1, emiss, &ann, &annp1, &bn);
    const double a1 = s2 * al[0], g1 = s2 * al[1];
    const double b1 = ((1.0 + a1) + g1) - g1 * bn;
    if (b1 == 0.0) { atomicOr(status, 1); return; }
    double Told = 0.0, Tc = Tat(1);
    double cprev = 0.0, bet = 0.0, uprev = 0.0;
    float bet32 = 0.0f, up32 = 0.0f;
    for (int j = 1; j <= nz; ++j) {
      const double Tn = j < nz ? Tat(j + 1) : 0.0;
      const double aj = j == 1 ? a1 : s2 * al[2 * (j - 1)];
      const double gj = j == 1 ? g1 : s2 * al[2 * (j - 1) + 1];
      double r, Arow, bj;
      if (j == 1) {
        r = (g1 * (annp1 + ann) + (((1.0 - a1) - g1) + g1 * bn) * Tc) + a1 * Tn;
        bj = b1;
        Arow = bj * Tc + (-a1) * Tn;
      } else if (j < nz) {
        r = (gj * Told + ((1.0 - aj) - gj) * Tc) + aj * Tn;
        bj = (1.0 + aj) + gj;
        Arow = ((-gj) * Told + bj * Tc) + (-aj) * Tn;
      } else {
        r = (gj * Told + (1.0 - gj) * Tc) + geo_term;
        bj = 1.0 + gj;
        Arow = (-gj) * Told + bj * Tc;
      }
      const double a_j = -gj;
      if (!MIXED) {
        double u;
        if (j == 1) {
          bet = bj;
          u = r / bet;
        } else {
          const double gm = cprev / bet;
          Gw[(int64_t)j * 64] = gm;
          bet = bj - a_j * gm;
          u = (r - a_j * uprev) / bet;
        }
        Tat(j) = u;  // old T_j is held in Tc
        uprev = u;
      } else {
        const float rho = (float)(r - Arow);
        float u;
        if (j == 1) {
          bet32 = (float)bj;
          u = rho * __builtin_amdgcn_rcpf(bet32);
        } else {
          const float gm = (float)cprev * __builtin_amdgcn_rcpf(bet32);
          G32[(int64_t)(2 * j) * nslot + slot] = gm;
          bet32 = (float)bj - (float)a_j * gm;
          u = (rho - (float)a_j * up32) * __builtin_amdgcn_rcpf(bet32);
        }
        G32[(int64_t)(2 * j + 1) * nslot + slot] = u;
        up32 = u;
      }
      cprev = -aj;
      Told = Tc;
      Tc = Tn;
    }
    double T1;
    if (!MIXED) {
      double unext = uprev;
      for (int j = nz - 1; j >= 1; --j) {
        const double u = Tat(j) - Gw[(int64_t)(j + 1) * 64] * unext;
        Tat(j) = u;
        unext = u;
      }
      T1 = unext;
    } else {
      float dnext = up32;
      Tat(nz) = Tat(nz) + (double)dnext;
      for (int j = nz - 1; j >= 1; --j) {
        const float d = G32[(int64_t)(2 * j + 1) * nslot + slot] -
                        G32[(int64_t)(2 * (j + 1)) * nslot + slot] * dnext;
        Tat(j) = Tat(j) + (double)d;
        dnext = d;
      }
      T1 = Tat(1);
    }
    const double Tsurf = 0.5 * ((annp1 + bn * T1) + T1);
    Fsurf = ((-kfirst) * (T1 - Tsurf)) / H.z1;
    Tat(0) = Tsurf;
    if (Tsurf_series) Tsurf_series[(int64_t)step * nf + f] = Tsurf;
    Qn = Qnp1;
  }
  Fs[f] = Fsurf;
  Qprev[f] = Qn;
}

// out[f * (nzmax + 1) + l] = T of facet f, layer l (NaN below its depth).
__global__ void gather_T_kernel(int64_t nf, int64_t nslot, int nzmax,
                                const int32_t *__restrict__ nzf, const int32_t *__restrict__ pos,
                                const double *__restrict__ T, double *__restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nf * (nzmax + 1)) return;
  const int64_t f = idx / (nzmax + 1);
  const int l = (int)(idx - f * (nzmax + 1));
  out[idx] = l <= nzf[f] ? T[(int64_t)l * nslot + pos[f]] : __builtin_nan("");
}

__global__ void gather_surface_kernel(int64_t nf, const int32_t *__restrict__ pos,
                                      const double *__restrict__ T, double *__restrict__ out) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f < nf) out[f] = T[pos[f]];
}

__global__ void scatter_T_kernel(int64_t nf, int64_t nslot, int nzmax,
                                 const int32_t *__restrict__ pos, const double *__restrict__ in,
                                 double *__restrict__ T) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nf * (nzmax + 1)) return;
  const int64_t f = idx / (nzmax + 1);
  const int l = (int)(idx - f * (nzmax + 1));
  T[(int64_t)l * nslot + pos[f]] = in[idx];
}

// out = -F (or out += -F): thermal_model::get_radiosity, thermal.hpp:61-63.
__global__ void neg_flux_kernel(int64_t nf, const double *__restrict__ F, double *__restrict__ out,
                                int accumulate) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  out[f] = accumulate ? out[f] - F[f] : -F[f];
}

__global__ void fill_kernel(int64_t n, double v, double *__restrict__ p) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

// ---------------------------------------------------------------------------
namespace {

// Host tables for a given dt, conductionQ.f:50-65 operation by operation.
// Register buckets (layer capacity of the cn_*_kernel<NZ> instantiations); 0
// means the streaming kernel.
constexpr int BUCKETS[] = {32, 48, 64, 80, 96, 128};
constexpr int NBUCKET = (int)(sizeof(BUCKETS) / sizeof(BUCKETS[0]));

int bucket_of(int nz) {
  for (int b : BUCKETS)
    if (nz <= b) return b;
  return 0;
}

void make_tables(const std::vector<int64_t> &goff, const double *z, const double *ti,
                 const double *rhoc, double dt, std::vector<double> &hdr,
                 std::vector<double> &coef, std::vector<float> &coefp,
                 std::vector<int64_t> &goffp) {
  const int ng = (int)goff.size() - 1;
  hdr.assign(6 * ng, 0.0);
  coef.assign(2 * goff[ng], 0.0);
  // FP32 table of the mixed kernel: 4 floats per row, each grid padded with
  // zero rows to its register bucket's depth
  goffp.assign(ng + 1, 0);
  for (int g = 0; g < ng; ++g) {
    const int nz = (int)(goff[g + 1] - goff[g]);
    const int nb = bucket_of(nz);
    goffp[g + 1] = goffp[g] + (nb ? nb : nz);
  }
  coefp.assign(4 * goffp[ng], 0.0f);
  for (int g = 0; g < ng; ++g) {
    const int64_t o = goff[g];
    const int nz = (int)(goff[g + 1] - o);
    const double *Z = z + o, *TI = ti + o, *RC = rhoc + o;
    std::vector<double> k(nz), alpha(nz, 0.0), gamma(nz, 0.0);
    for (int i = 0; i < nz; ++i) k[i] = (TI[i] * TI[i]) / RC[i];
    const double dz = 2.0 * Z[0];
    const double beta = (dt / RC[0]) / (2.0 * (dz * dz));
    alpha[0] = beta * k[1];
    gamma[0] = beta * k[0];
    for (int i = 1; i < nz - 1; ++i) {
      const double buf = dt / (Z[i + 1] - Z[i - 1]);
      alpha[i] = (((2.0 * k[i + 1]) * buf) / (RC[i] + RC[i + 1])) / (Z[i + 1] - Z[i]);
      gamma[i] = (((2.0 * k[i]) * buf) / (RC[i] + RC[i + 1])) / (Z[i] - Z[i - 1]);
    }
    const double dl = Z[nz - 1] - Z[nz - 2];
    const double buf = dt / (dl * dl);
    gamma[nz - 1] = (k[nz - 1] * buf) / (2.0 * RC[nz - 1]);
    for (int i = 0; i < nz; ++i) {
      coef[2 * (o + i)] = alpha[i];
      coef[2 * (o + i) + 1] = gamma[i];
      float *cp = &coefp[4 * (goffp[g] + i)];
      cp[0] = (float)alpha[i];  // 0 on the bottom row (no row below it)
      cp[1] = (float)gamma[i];
      cp[2] = (float)(alpha[i] + gamma[i]);
      cp[3] = i == nz - 1 ? 1.0f : 0.0f;
    }
    double *h = &hdr[6 * g];
    h[0] = k[0] / dz;
    h[1] = k[0];
    h[2] = Z[0];
    h[3] = dt / RC[nz - 1];
    h[4] = dl;
    h[5] = nz;
  }
}


}  // namespace

struct ThermalImpl {
  std::vector<double> z, ti, rhoc;
  double dt = -1;
  double *d_hdr = nullptr, *d_coef = nullptr;
  float *d_coefp = nullptr;
  int64_t *d_goffp = nullptr;
  ~ThermalImpl() { dfree(d_hdr); dfree(d_coef); dfree(d_coefp); dfree(d_goffp); }
};

static ThermalImpl *impl_of(illum_thermal *th) { return (ThermalImpl *)th->impl; }

illum_thermal *thermal_create(illum_ctx *ctx, int64_t nf, int ngrid, const int64_t *grid_off,
                              const double *z, const double *ti, const double *rhoc,
                              const int32_t *grid_of_facet, const double *ti_scale,
                              double emissivity, double T0, const double *T0_full,
                              int precision) {
  if (nf <= 0 || ngrid <= 0 || !grid_off || !z || !ti || !rhoc)
    throw Error(ILLUM_ERR_ARG, "thermal model needs facets and at least one layer grid");
  if (nf >= ((int64_t)1 << 31)) throw Error(ILLUM_ERR_CAPACITY, "too many facets");
  if (precision != ILLUM_THERM_FP64 && precision != ILLUM_THERM_MIXED)
    throw Error(ILLUM_ERR_ARG, "unknown thermal precision");
  if (grid_off[0] != 0) throw Error(ILLUM_ERR_ARG, "grid_off[0] must be 0");
  std::vector<int64_t> goff(grid_off, grid_off + ngrid + 1);
  int nzmax = 0;
  for (int g = 0; g < ngrid; ++g) {
    const int64_t nz = goff[g + 1] - goff[g];
    if (nz < 2) throw Error(ILLUM_ERR_ARG, "every layer grid needs at least 2 layers");
    if (nz > 1000) throw Error(ILLUM_ERR_CAPACITY, "at most 1000 layers (conductionQ.f:38 NMAX)");
    for (int64_t i = goff[g]; i < goff[g + 1]; ++i) {
      if (!(z[i] > 0) || (i > goff[g] && !(z[i] > z[i - 1])))
        throw Error(ILLUM_ERR_ARG, "layer depths must be positive and increasing");
      if (!(rhoc[i] > 0) || !(ti[i] > 0)) throw Error(ILLUM_ERR_ARG, "ti and rhoc must be positive");
    }
    nzmax = std::max<int>(nzmax, (int)nz);
  }
  std::vector<int32_t> nzh(nf), gof(nf);
  std::vector<double> s2(nf, 1.0);
  std::vector<std::vector<int32_t>> members(ngrid);
  for (int64_t f = 0; f < nf; ++f) {
    const int g = grid_of_facet ? grid_of_facet[f] : 0;
    if (g < 0 || g >= ngrid) throw Error(ILLUM_ERR_ARG, "grid_of_facet out of range");
    gof[f] = g;
    members[g].push_back((int32_t)f);
    nzh[f] = (int32_t)(goff[g + 1] - goff[g]);
    if (ti_scale) {
      if (!(ti_scale[f] > 0)) throw Error(ILLUM_ERR_ARG, "ti_scale must be positive");
      s2[f] = ti_scale[f] * ti_scale[f];
    }
  }
  std::unique_ptr<illum_thermal, void (*)(illum_thermal *)> th(new illum_thermal(), destroy_thermal);
  th->ctx = ctx;
  th->nf = nf;
  th->nzmax = nzmax;
  th->precision = precision;
  th->ngrid = ngrid;
  th->emiss = emissivity;
  th->goff = goff;
  th->nz_host = nzh;

  // Slot permutation: grids grouped by layer bucket, each grid's facets
  // padded to whole wavefronts.
  std::vector<int32_t> fidx, wgrid, pos(nf, -1);
  bool need_stream = false;
  for (int bi = 0; bi <= NBUCKET; ++bi) {
    const int bucket = bi < NBUCKET ? BUCKETS[bi] : 0;
    const int64_t block0 = (int64_t)wgrid.size();
    for (int g = 0; g < ngrid; ++g) {
      if (members[g].empty() || bucket_of((int)(goff[g + 1] - goff[g])) != bucket) continue;
      for (size_t k = 0; k < members[g].size(); k += 64) {
        for (size_t q = k; q < k + 64; ++q) {
          if (q < members[g].size()) {
            pos[members[g][q]] = (int32_t)fidx.size();
            fidx.push_back(members[g][q]);
          } else {
            fidx.push_back(-1);
          }
        }
        wgrid.push_back(g);
      }
    }
    const int64_t nb = (int64_t)wgrid.size() - block0;
    if (nb > 0) {
      th->launches.push_back({bucket, block0, nb});
      if (bucket == 0) need_stream = true;
    }
  }
  const int64_t nslot = (int64_t)fidx.size();
  th->nslot = nslot;
  hipStream_t s = stream_of(ctx);
  auto up = [&](auto *&dst, const auto &v) {
    using T = typename std::remove_reference<decltype(*dst)>::type;
    dst = dalloc<T>(v.size());
    ILLUM_HIP(hipMemcpyAsync(dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
  };
  up(th->d_fidx, fidx);
  up(th->d_pos, pos);
  up(th->d_wgrid, wgrid);
  up(th->d_nz, nzh);
  up(th->d_grid, gof);
  if (ti_scale) up(th->d_s2, s2);
  up(th->d_goff, goff);
  th->d_T = dalloc<double>((size_t)(nzmax + 1) * nslot);
  th->d_Qprev = dalloc<double>(nf);
  th->d_F = dalloc<double>(nf);
  // Thomas scratch rows (every engine streams its multipliers through them)
  // Scratch rows for the engines that stream Thomas multipliers: the
  // streaming kernel (either precision, or forced for comparison) and the
  // FP64 register kernel above 96 layers.
  if (precision == ILLUM_THERM_FP64) th->d_work = dalloc<double>((size_t)(nzmax + 1) * nslot);
  else th->d_work32 = dalloc<float>((size_t)(2 * nzmax + 2) * nslot);
  (void)need_stream;
  th->d_status = dalloc<int>(1);
  ILLUM_HIP(hipMemsetAsync(th->d_Qprev, 0, 8 * nf, s));
  ILLUM_HIP(hipMemsetAsync(th->d_F, 0, 8 * nf, s));
  ILLUM_HIP(hipMemsetAsync(th->d_status, 0, 4, s));
  {
    const int64_t tot = (int64_t)(nzmax + 1) * nslot;
    fill_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(tot, T0, th->d_T);
    ILLUM_LAUNCH_CHECK();
  }
  if (T0_full) {
    DBuf<double> tmp((size_t)nf * (nzmax + 1));
    ILLUM_HIP(hipMemcpyAsync(tmp.p, T0_full, 8 * nf * (nzmax + 1), hipMemcpyHostToDevice, s));
    const int64_t tot = nf * (nzmax + 1);
    scatter_T_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(nf, nslot, nzmax, th->d_pos,
                                                                 tmp.p, th->d_T);
    ILLUM_LAUNCH_CHECK();
    sync(ctx);
  }
  auto *im = new ThermalImpl();
  im->z.assign(z, z + goff[ngrid]);
  im->ti.assign(ti, ti + goff[ngrid]);
  im->rhoc.assign(rhoc, rhoc + goff[ngrid]);
  th->impl = im;
  sync(ctx);
  return th.release();
}

void thermal_free_impl(illum_thermal *th) {
  delete impl_of(th);
  th->impl = nullptr;
}

static void ensure_tables(illum_thermal *th, double dt) {
  ThermalImpl *im = impl_of(th);
  if (im->dt == dt && im->d_hdr) return;
  std::vector<double> hdr, coef;
  std::vector<float> coefp;
  std::vector<int64_t> goffp;
  make_tables(th->goff, im->z.data(), im->ti.data(), im->rhoc.data(), dt, hdr, coef, coefp,
              goffp);
  dfree(im->d_hdr);
  dfree(im->d_coef);
  dfree(im->d_coefp);
  dfree(im->d_goffp);
  im->d_hdr = dalloc<double>(hdr.size());
  im->d_coef = dalloc<double>(coef.size());
  im->d_coefp = dalloc<float>(coefp.size());
  im->d_goffp = dalloc<int64_t>(goffp.size());
  hipStream_t s = stream_of(th->ctx);
  ILLUM_HIP(hipMemcpyAsync(im->d_hdr, hdr.data(), 8 * hdr.size(), hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(im->d_coef, coef.data(), 8 * coef.size(), hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(im->d_coefp, coefp.data(), 4 * coefp.size(), hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(im->d_goffp, goffp.data(), 8 * goffp.size(), hipMemcpyHostToDevice, s));
  sync(th->ctx);
  im->dt = dt;
}

// Form of the mixed register kernel: 0 = cn_mixed_kernel (Q, R per lane, no
// coefficient loads in the step loop), 1 = cn_mixed_p_kernel (pivots per
// lane, coefficients reloaded).  Form 0 is the faster in every bucket,
// including 80 and 96 layers where its arrays spill a few registers (327,680
// facets x 4,096 steps, 32-96 layers: 57 ms against 79 ms).
// ILLUM_CN_MIXED_FORM=1 selects form 1, for comparison.
static int cn_mixed_form(int) {
  const char *e = getenv("ILLUM_CN_MIXED_FORM");
  return e && atoi(e) == 1 ? 1 : 0;
}

template <bool MIXED>
static void launch_all(illum_thermal *th, ThermalImpl *im, hipStream_t s0, double Fgeo,
                       int nsteps, int64_t qs, const double *dQ, double *Tss,
                       bool fork_ok = true) {
  const char *eng = getenv("ILLUM_CN_ENGINE");
  const bool force_stream = eng && std::string(eng) == "stream";
  // The depth buckets touch disjoint slots and facets.  At one or two
  // wavefronts per SIMD a bucket's launch ends with a partly idle last round,
  // so with several buckets they run concurrently on up to three forked
  // streams (largest first), each filling the others' tails.
  const int nl = (int)th->launches.size();
  const int nfork = nl > 1 && fork_ok ? std::min(nl, 3) : 0;
  std::vector<hipStream_t> fs;
  if (nfork) fs = fork_streams(th->ctx, nfork);
  std::vector<int> order(nl);
  for (int i = 0; i < nl; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    const auto &A = th->launches[a], &B = th->launches[b];
    return A.nblocks * std::max(A.nz_bucket, 128) > B.nblocks * std::max(B.nz_bucket, 128);
  });
  for (int oi = 0; oi < nl; ++oi) {
    const auto &L = th->launches[order[oi]];
    const hipStream_t s = nfork ? fs[oi % nfork] : s0;
#define ILLUM_CN_ARGS                                                                        \
  th->d_fidx, th->d_wgrid, L.block0, th->nf, th->nslot, th->d_s2, im->d_hdr, im->d_coef,     \
      th->d_goff, th->emiss, Fgeo, nsteps, qs, dQ, th->d_Qprev, th->d_T, th->d_F
    const unsigned nb = (unsigned)L.nblocks;
    switch (force_stream ? 0 : L.nz_bucket) {
#define ILLUM_CN_REG(NZ, GMREG)                                                              \
  case NZ:                                                                                   \
    if (MIXED)                                                                               \
      (cn_mixed_form(NZ) == 1 ? cn_mixed_p_kernel<NZ>                                         \
       : Fgeo != 0.0          ? cn_mixed_kernel<NZ, true>                                     \
                              : cn_mixed_kernel<NZ, false>)<<<nb, 64, 0, s>>>(                \
          th->d_fidx, th->d_wgrid, L.block0, th->nf, th->nslot, th->d_s2, im->d_hdr,         \
          im->d_coef, (const float4 *)im->d_coefp, th->d_goff, im->d_goffp, th->emiss, Fgeo, \
          nsteps, qs, dQ, th->d_Qprev, th->d_T, th->d_F, Tss);                                   \
    else                                                                                     \
      cn_fp64_kernel<NZ, GMREG><<<nb, 64, 0, s>>>(ILLUM_CN_ARGS, th->d_work, Tss,            \
                                                  th->d_status, th->nzmax + 1);              \
    break;
      ILLUM_CN_REG(32, true)
      ILLUM_CN_REG(48, true)
      ILLUM_CN_REG(64, true)
      ILLUM_CN_REG(80, true)
      ILLUM_CN_REG(96, true)
      ILLUM_CN_REG(128, false)
#undef ILLUM_CN_REG
      default:
        cn_stream_kernel<MIXED><<<nb, 64, 0, s>>>(ILLUM_CN_ARGS, th->d_work, th->d_work32, Tss,
                                                  th->d_status, th->nzmax + 1);
    }
#undef ILLUM_CN_ARGS
    ILLUM_LAUNCH_CHECK();
  }
  if (nfork) join_streams(th->ctx, nfork);
}

static void check_status(illum_thermal *th) {
  int st = 0;
  ILLUM_HIP(hipMemcpyAsync(&st, th->d_status, 4, hipMemcpyDeviceToHost, stream_of(th->ctx)));
  sync(th->ctx);
  if (st) throw Error(ILLUM_ERR_ZERO_PIVOT, "tridag: rewrite equations (b(1) == 0)");
}

void thermal_run(illum_thermal *th, double dt, const double *Q, int nsteps, int flags,
                 double Fgeo, double *Tsurf_series) {
  if (nsteps < 1) throw Error(ILLUM_ERR_ARG, "nsteps must be >= 1");
  if (!(dt > 0)) throw Error(ILLUM_ERR_ARG, "dt must be positive");
  if (!Q) throw Error(ILLUM_ERR_ARG, "null Q");
  illum_ctx *ctx = th->ctx;
  hipStream_t s = stream_of(ctx);
  ensure_tables(th, dt);
  ThermalImpl *im = impl_of(th);
  const int64_t nf = th->nf;
  // ILLUM_Q_HELD: Q is one row of nf values, the flux of every step (a row
  // stride of 0), the same steps as nsteps single steps with that Q
  const bool held = flags & ILLUM_Q_HELD;
  const int64_t qs = held ? 0 : nf;
  const double *dQ = Q;
  if (!(flags & ILLUM_IN_ON_DEVICE)) {
    const int64_t need = held ? nf : nf * nsteps;
    if (th->qdev_cap < need) {
      dfree(th->d_Qdev);
      th->d_Qdev = dalloc<double>(need);
      th->qdev_cap = need;
    }
    ILLUM_HIP(hipMemcpyAsync(th->d_Qdev, Q, 8 * need, hipMemcpyHostToDevice, s));
    dQ = th->d_Qdev;
  }
  const bool mixed = th->precision == ILLUM_THERM_MIXED;
  auto launch = [&](bool fork) {
    if (mixed) launch_all<true>(th, im, s, Fgeo, nsteps, qs, dQ, Tsurf_series, fork);
    else launch_all<false>(th, im, s, Fgeo, nsteps, qs, dQ, Tsurf_series, fork);
  };
  // With several depth buckets, whether they run concurrently on forked
  // streams or one after another does not change the results (they touch
  // disjoint slots); the autotuner times both once per model shape, on a
  // snapshot of the state restored before each trial and after tuning.
  bool fork = true;
  if (th->launches.size() > 1) {
    std::string key = std::string("cn_streams[") + (mixed ? "mixed" : "fp64") +
                      ",nslot=" + std::to_string(th->nslot) + ",buckets=" +
                      std::to_string(th->launches.size()) + ",steps=" + std::to_string(nsteps) + "]";
    const bool cached = ctx->tune_cache.count(key) != 0;
    const char *off = getenv("ILLUM_NO_AUTOTUNE");
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    ILLUM_HIP(hipStreamIsCapturing(s, &cs));
    if (cached || (off && off[0] == '1') || cs != hipStreamCaptureStatusNone) {
      fork = autotune(ctx, key, {1, 0}, 1, [&](int) {}, false) == 1;
    } else {
      const size_t tb = 8 * (size_t)(th->nzmax + 1) * th->nslot;
      DBuf<double> snap(tb / 8 + 2 * th->nf);
      auto copy = [&](bool save) {
        double *dT = snap.p, *dq = snap.p + tb / 8, *dF = dq + th->nf;
        auto cp = [&](void *a, void *b, size_t n) {
          ILLUM_HIP(hipMemcpyAsync(save ? a : b, save ? b : a, n, hipMemcpyDeviceToDevice, s));
        };
        cp(dT, th->d_T, tb);
        cp(dq, th->d_Qprev, 8 * th->nf);
        cp(dF, th->d_F, 8 * th->nf);
      };
      copy(true);
      fork = autotune(ctx, key, {1, 0}, 1, [&](int c) { copy(false); launch(c == 1); }) == 1;
      copy(false);
      sync(ctx);
    }
  }
  {
    StageTimer st(ctx, mixed ? "cn_mixed" : "cn_fp64");
    launch(fork);
  }
  // A zero pivot is sticky in d_status; an asynchronous step leaves its check
  // to the next synchronous step or thermal_get.
  if ((flags & ILLUM_IN_ON_DEVICE) && (flags & ILLUM_ASYNC)) return;
  check_status(th);
}

// Prepares the coefficient tables of step dt (host work, may allocate), so
// that thermal_enqueue can then be captured.
void thermal_prepare(illum_thermal *th, double dt) {
  if (!(dt > 0)) throw Error(ILLUM_ERR_ARG, "dt must be positive");
  ensure_tables(th, dt);
}

// Queues nsteps steps with the device Q series dQ on the context stream, on
// that stream alone (no forked streams, no status readback, no host work):
// graph-capturable once thermal_prepare(th, dt) has run.  A zero pivot stays
// in the model's status word for the next checked call.
void thermal_enqueue(illum_thermal *th, double dt, const double *dQ, int nsteps) {
  ThermalImpl *im = impl_of(th);
  if (im->dt != dt || !im->d_hdr) throw Error(ILLUM_ERR_ARG, "thermal tables not prepared for dt");
  hipStream_t s = stream_of(th->ctx);
  if (th->precision == ILLUM_THERM_MIXED) launch_all<true>(th, im, s, 0.0, nsteps, th->nf, dQ, nullptr, false);
  else launch_all<false>(th, im, s, 0.0, nsteps, th->nf, dQ, nullptr, false);
}

// The model's sticky status: ILLUM_ERR_ZERO_PIVOT after a zero pivot.
void thermal_check(illum_thermal *th) { check_status(th); }

void thermal_get(illum_thermal *th, double *T, double *Tsurf, double *Fsurf, int flags) {
  illum_ctx *ctx = th->ctx;
  hipStream_t s = stream_of(ctx);
  const int64_t nf = th->nf;
  const bool dev = flags & ILLUM_OUT_ON_DEVICE;
  if (T) {
    const int64_t tot = nf * (th->nzmax + 1);
    DBuf<double> tmp;
    double *dst = T;
    if (!dev) { tmp.reset(tot); dst = tmp.p; }
    gather_T_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(
        nf, th->nslot, th->nzmax, th->d_nz, th->d_pos, th->d_T, dst);
    ILLUM_LAUNCH_CHECK();
    if (!dev) ILLUM_HIP(hipMemcpyAsync(T, dst, 8 * tot, hipMemcpyDeviceToHost, s));
    sync(ctx);
  }
  if (Tsurf) {
    DBuf<double> tmp;
    double *dst = Tsurf;
    if (!dev) { tmp.reset(nf); dst = tmp.p; }
    gather_surface_kernel<<<(unsigned)((nf + 255) / 256), 256, 0, s>>>(nf, th->d_pos, th->d_T, dst);
    ILLUM_LAUNCH_CHECK();
    if (!dev) ILLUM_HIP(hipMemcpyAsync(Tsurf, dst, 8 * nf, hipMemcpyDeviceToHost, s));
    sync(ctx);
  }
  if (Fsurf)
    ILLUM_HIP(hipMemcpyAsync(Fsurf, th->d_F, 8 * nf, dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  check_status(th);
}

void thermal_get_radiosity(illum_thermal *th, double *out, int flags) {
  illum_ctx *ctx = th->ctx;
  hipStream_t s = stream_of(ctx);
  const int64_t nf = th->nf;
  const bool dev = flags & ILLUM_OUT_ON_DEVICE;
  const bool acc = flags & ILLUM_ACCUMULATE;
  DBuf<double> tmp;
  double *dst = out;
  if (!dev) {
    tmp.reset(nf);
    dst = tmp.p;
    if (acc) ILLUM_HIP(hipMemcpyAsync(dst, out, 8 * nf, hipMemcpyHostToDevice, s));
  }
  neg_flux_kernel<<<(unsigned)((nf + 255) / 256), 256, 0, s>>>(nf, th->d_F, dst, acc ? 1 : 0);
  ILLUM_LAUNCH_CHECK();
  if (!dev) {
    ILLUM_HIP(hipMemcpyAsync(out, dst, 8 * nf, hipMemcpyDeviceToHost, s));
    sync(ctx);
  }
}

}  // namespace illum
