// Batched per-facet Crank-Nicolson subsurface conduction
// (thermal_model::step, thermal.hpp:65-105 -> conductionQ, conductionQ.f:35-98,
// tridiagonal solve per tridag.for:17-27, i.e. the Thomas algorithm; see
// Golub & Van Loan, Matrix Computations, section 4.3).
//
// Layout: facets are independent; one lane owns one facet for all steps of a
// launch.  Temperatures live in HBM layer-major, T[l * nfp + f] (l = 0 is the
// surface, l = 1..nz the layers), so a wavefront touches 64 consecutive
// facets of one layer per access (fully coalesced).  Ragged layer counts use
// the same stride; a facet with fewer layers simply stops early.
//
// Coefficients: the time-invariant tridiagonal entries of each layer grid are
// tabulated on the host in conductionQ's own operation order for the current
// dt; a facet's thermal-inertia scale s multiplies every conductivity by s^2.
//
// FP64 mode: the reference step verbatim -- right-hand side, Thomas sweep
// with one division per row, surface temperature and flux, all in FP64.
// Mixed mode: iterative refinement from the previous state.  The FP64
// residual rho = r - A T^n is formed in the forward pass, the correction
// A d = rho is solved by an FP32 Thomas sweep (hardware reciprocal), and
// T^{n+1} = T^n + d in FP64.  Since |d| is the per-step temperature change,
// the FP32 error is relative to that change rather than to T.  The radiative
// surface linearisation and surface flux stay FP64.
#include "hip_util.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace illum {

constexpr double SIGMA_SB = 5.6704e-8;  // conductionQ.f:39

// Per-grid table entry for layer i (0-based): alpha, gamma.
// Per-grid header (6 doubles): k1 = k(1)/dz, k(1), z(1), dt/rhoc(nz),
// z(nz) - z(nz-1), nz.
struct GridHdr {
  double k1, kfirst, z1, geo, dzlast, nz;
};

template <bool MIXED>
__global__ __launch_bounds__(256) void cn_kernel(
    int64_t nf, int64_t nfp, const int32_t *__restrict__ nzf, const int32_t *__restrict__ gridf,
    const double *__restrict__ s2f, const double *__restrict__ hdr,
    const double *__restrict__ coef, const int64_t *__restrict__ goff, double emiss,
    double Fgeo, int nsteps, const double *__restrict__ Q, const double *__restrict__ Qprev,
    double *__restrict__ T, double *__restrict__ Fs, double *__restrict__ G,
    float *__restrict__ G32, double *__restrict__ Tsurf_series, int *__restrict__ status) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int nz = nzf[f];
  const int g = gridf ? gridf[f] : 0;
  const double s2 = s2f ? s2f[f] : 1.0;
  const GridHdr H = *(const GridHdr *)(hdr + 6 * g);
  const double *al = coef + 2 * goff[g];  // interleaved alpha, gamma
  const double k1 = s2 * H.k1, kfirst = s2 * H.kfirst;
  const double geo_term = ((H.geo) * Fgeo) / H.dzlast;
  float *gam32 = G32;  // row 2j: FP32 multiplier gam_j, row 2j+1: FP32 forward value u_j
  double Tsurf = T[f];
  double Qn = Qprev[f];
  double Fsurf = Fs[f];
  for (int step = 0; step < nsteps; ++step) {
    const double Qnp1 = Q[(int64_t)step * nf + f];
    // Radiative boundary linearised at the old surface temperature.
    const double Tr = Tsurf;
    const double Tr3 = (Tr * Tr) * Tr, Tr4 = Tr3 * Tr;  // Tr**3, Tr**4 left to right
    const double arad = ((-3.0 * emiss) * SIGMA_SB) * Tr4;
    const double brad = ((2.0 * emiss) * SIGMA_SB) * Tr3;
    const double den = k1 + brad;
    const double ann = (Qn - arad) / den;
    const double annp1 = (Qnp1 - arad) / den;
    const double bn = (k1 - brad) / den;
    const double a1 = s2 * al[0], g1 = s2 * al[1];
    const double b1 = ((1.0 + a1) + g1) - g1 * bn;
    if (b1 == 0.0) { atomicOr(status, 1); return; }

    // Forward sweep over layers 1..nz, right-hand side formed on the fly
    // from the old temperatures.
    double Tm = 0.0;                 // old T_{j-1}
    double Tc = T[1 * nfp + f];      // old T_j
    double aj = a1, gj = g1;         // alpha_j, gamma_j
    double bj = b1;
    double cprev = 0.0;              // c_{j-1} = -alpha_{j-1}
    double bet = 0.0;
    double uprev = 0.0;
    float bet32 = 0.0f, up32 = 0.0f;
    for (int j = 1; j <= nz; ++j) {
      const double Tn = j < nz ? T[(int64_t)(j + 1) * nfp + f] : 0.0;
      double r, Arow;
      if (j == 1) {
        r = (g1 * (annp1 + ann) + (((1.0 - a1) - g1) + g1 * bn) * Tc) + a1 * Tn;
        Arow = b1 * Tc + (-a1) * Tn;
      } else {
        aj = s2 * al[2 * (j - 1)];
        gj = s2 * al[2 * (j - 1) + 1];
        if (j < nz) {
          r = (gj * Tm + ((1.0 - aj) - gj) * Tc) + aj * Tn;
          bj = (1.0 + aj) + gj;
          Arow = ((-gj) * Tm + bj * Tc) + (-aj) * Tn;
        } else {
          r = (gj * Tm + (1.0 - gj) * Tc) + geo_term;
          bj = 1.0 + gj;
          Arow = (-gj) * Tm + bj * Tc;
        }
      }
      const double a_j = -gj;  // sub-diagonal (unused for j == 1)
      if (!MIXED) {
        double u;
        if (j == 1) {
          bet = bj;
          u = r / bet;
        } else {
          const double gam = cprev / bet;
          G[(int64_t)j * nfp + f] = gam;
          bet = bj - a_j * gam;
          u = (r - a_j * uprev) / bet;
        }
        T[(int64_t)j * nfp + f] = u;  // old T_j is held in Tc
        uprev = u;
      } else {
        const float rho = (float)(r - Arow);
        float u;
        if (j == 1) {
          bet32 = (float)bj;
          u = rho * __builtin_amdgcn_rcpf(bet32);
        } else {
          const float gam = (float)cprev * __builtin_amdgcn_rcpf(bet32);
          gam32[(int64_t)(2 * j) * nfp + f] = gam;
          bet32 = (float)bj - (float)a_j * gam;
          u = (rho - (float)a_j * up32) * __builtin_amdgcn_rcpf(bet32);
        }
        gam32[(int64_t)(2 * j + 1) * nfp + f] = u;
        up32 = u;
      }
      cprev = -aj;
      Tm = Tc;
      Tc = Tn;
    }
    // Back substitution.
    double T1;
    if (!MIXED) {
      double unext = uprev;  // u_nz, already stored
      for (int j = nz - 1; j >= 1; --j) {
        const double u = T[(int64_t)j * nfp + f] - G[(int64_t)(j + 1) * nfp + f] * unext;
        T[(int64_t)j * nfp + f] = u;
        unext = u;
      }
      T1 = unext;
    } else {
      float dnext = up32;
      T[(int64_t)nz * nfp + f] = T[(int64_t)nz * nfp + f] + (double)dnext;
      for (int j = nz - 1; j >= 1; --j) {
        const float d = gam32[(int64_t)(2 * j + 1) * nfp + f] -
                        gam32[(int64_t)(2 * (j + 1)) * nfp + f] * dnext;
        T[(int64_t)j * nfp + f] = T[(int64_t)j * nfp + f] + (double)d;
        dnext = d;
      }
      T1 = T[1 * nfp + f];
    }
    Tsurf = 0.5 * ((annp1 + bn * T1) + T1);
    Fsurf = ((-kfirst) * (T1 - Tsurf)) / H.z1;
    T[f] = Tsurf;
    if (Tsurf_series) Tsurf_series[(int64_t)step * nf + f] = Tsurf;
    Qn = Qnp1;
  }
  Fs[f] = Fsurf;
}

__global__ void copy_qprev_kernel(int64_t nf, const double *__restrict__ Qlast,
                                  double *__restrict__ Qprev) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f < nf) Qprev[f] = Qlast[f];
}

__global__ void gather_T_kernel(int64_t nf, int64_t nfp, int nzmax, const int32_t *__restrict__ nzf,
                                const double *__restrict__ T, double *__restrict__ out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nf * (nzmax + 1)) return;
  const int64_t f = idx / (nzmax + 1);
  const int l = (int)(idx - f * (nzmax + 1));
  out[idx] = l <= nzf[f] ? T[(int64_t)l * nfp + f] : __builtin_nan("");
}

__global__ void scatter_T_kernel(int64_t nf, int64_t nfp, int nzmax, const double *__restrict__ in,
                                 double *__restrict__ T) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nf * (nzmax + 1)) return;
  const int64_t f = idx / (nzmax + 1);
  const int l = (int)(idx - f * (nzmax + 1));
  T[(int64_t)l * nfp + f] = in[idx];
}

__global__ void fill_kernel(int64_t n, double v, double *__restrict__ p) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

// ---------------------------------------------------------------------------
namespace {

// Host tables for a given dt, conductionQ.f:50-65 operation by operation.
void make_tables(const std::vector<int64_t> &goff, const double *z, const double *ti,
                 const double *rhoc, double dt, std::vector<double> &hdr,
                 std::vector<double> &coef) {
  const int ng = (int)goff.size() - 1;
  hdr.assign(6 * ng, 0.0);
  coef.assign(2 * goff[ng], 0.0);
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
  ~ThermalImpl() { dfree(d_hdr); dfree(d_coef); }
};

static ThermalImpl *impl_of(illum_thermal *th);
illum_thermal *thermal_create(illum_ctx *ctx, int64_t nf, int ngrid, const int64_t *grid_off,
                              const double *z, const double *ti, const double *rhoc,
                              const int32_t *grid_of_facet, const double *ti_scale,
                              double emissivity, double T0, const double *T0_full,
                              int precision) {
  if (nf <= 0 || ngrid <= 0 || !grid_off || !z || !ti || !rhoc)
    throw Error(ILLUM_ERR_ARG, "thermal model needs facets and at least one layer grid");
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
  std::vector<int32_t> nzh(nf);
  std::vector<double> s2(nf, 1.0);
  for (int64_t f = 0; f < nf; ++f) {
    const int g = grid_of_facet ? grid_of_facet[f] : 0;
    if (g < 0 || g >= ngrid) throw Error(ILLUM_ERR_ARG, "grid_of_facet out of range");
    nzh[f] = (int32_t)(goff[g + 1] - goff[g]);
    if (ti_scale) {
      if (!(ti_scale[f] > 0)) throw Error(ILLUM_ERR_ARG, "ti_scale must be positive");
      s2[f] = ti_scale[f] * ti_scale[f];
    }
  }
  std::unique_ptr<illum_thermal, void (*)(illum_thermal *)> th(new illum_thermal(), destroy_thermal);
  th->ctx = ctx;
  th->nf = nf;
  th->nfp = (nf + 63) / 64 * 64;
  th->nzmax = nzmax;
  th->precision = precision;
  th->ngrid = ngrid;
  th->emiss = emissivity;
  th->goff = goff;
  th->nz_host = nzh;
  hipStream_t s = stream_of(ctx);
  const int64_t nfp = th->nfp;
  th->d_nz = dalloc<int32_t>(nf);
  ILLUM_HIP(hipMemcpyAsync(th->d_nz, nzh.data(), 4 * nf, hipMemcpyHostToDevice, s));
  if (grid_of_facet) {
    th->d_grid = dalloc<int32_t>(nf);
    ILLUM_HIP(hipMemcpyAsync(th->d_grid, grid_of_facet, 4 * nf, hipMemcpyHostToDevice, s));
  }
  if (ti_scale) {
    th->d_s2 = dalloc<double>(nf);
    ILLUM_HIP(hipMemcpyAsync(th->d_s2, s2.data(), 8 * nf, hipMemcpyHostToDevice, s));
  }
  th->d_goff = dalloc<int64_t>(ngrid + 1);
  ILLUM_HIP(hipMemcpyAsync(th->d_goff, goff.data(), 8 * (ngrid + 1), hipMemcpyHostToDevice, s));
  th->d_T = dalloc<double>((size_t)(nzmax + 1) * nfp);
  th->d_Qprev = dalloc<double>(nf);
  th->d_F = dalloc<double>(nf);
  if (precision == ILLUM_THERM_FP64) th->d_work = dalloc<double>((size_t)(nzmax + 1) * nfp);
  else th->d_work32 = dalloc<float>((size_t)(2 * nzmax + 2) * nfp);
  th->d_status = dalloc<int>(1);
  ILLUM_HIP(hipMemsetAsync(th->d_Qprev, 0, 8 * nf, s));
  ILLUM_HIP(hipMemsetAsync(th->d_F, 0, 8 * nf, s));
  ILLUM_HIP(hipMemsetAsync(th->d_status, 0, 4, s));
  if (T0_full) {
    DBuf<double> tmp((size_t)nf * (nzmax + 1));
    ILLUM_HIP(hipMemcpyAsync(tmp.p, T0_full, 8 * nf * (nzmax + 1), hipMemcpyHostToDevice, s));
    const int64_t tot = nf * (nzmax + 1);
    scatter_T_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(nf, nfp, nzmax, tmp.p, th->d_T);
    ILLUM_LAUNCH_CHECK();
    sync(ctx);
  } else {
    const int64_t tot = (int64_t)(nzmax + 1) * nfp;
    fill_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(tot, T0, th->d_T);
    ILLUM_LAUNCH_CHECK();
  }
  auto *im = new ThermalImpl();
  im->z.assign(z, z + goff[ngrid]);
  im->ti.assign(ti, ti + goff[ngrid]);
  im->rhoc.assign(rhoc, rhoc + goff[ngrid]);
  th->impl = im;
  sync(ctx);
  return th.release();
}

static ThermalImpl *impl_of(illum_thermal *th) { return (ThermalImpl *)th->impl; }

void thermal_free_impl(illum_thermal *th) {
  delete impl_of(th);
  th->impl = nullptr;
}

static void ensure_tables(illum_thermal *th, double dt) {
  ThermalImpl *im = impl_of(th);
  if (im->dt == dt && im->d_hdr) return;
  std::vector<double> hdr, coef;
  make_tables(th->goff, im->z.data(), im->ti.data(), im->rhoc.data(), dt, hdr, coef);
  dfree(im->d_hdr);
  dfree(im->d_coef);
  im->d_hdr = dalloc<double>(hdr.size());
  im->d_coef = dalloc<double>(coef.size());
  hipStream_t s = stream_of(th->ctx);
  ILLUM_HIP(hipMemcpyAsync(im->d_hdr, hdr.data(), 8 * hdr.size(), hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(im->d_coef, coef.data(), 8 * coef.size(), hipMemcpyHostToDevice, s));
  sync(th->ctx);
  im->dt = dt;
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
  const double *dQ = Q;
  if (!(flags & ILLUM_IN_ON_DEVICE)) {
    const int64_t need = nf * nsteps;
    if (th->qdev_cap < need) {
      dfree(th->d_Qdev);
      th->d_Qdev = dalloc<double>(need);
      th->qdev_cap = need;
    }
    ILLUM_HIP(hipMemcpyAsync(th->d_Qdev, Q, 8 * need, hipMemcpyHostToDevice, s));
    dQ = th->d_Qdev;
  }
  const unsigned blocks = (unsigned)((nf + 255) / 256);
  {
    StageTimer st(ctx, th->precision == ILLUM_THERM_MIXED ? "cn_mixed" : "cn_fp64");
    if (th->precision == ILLUM_THERM_MIXED)
      cn_kernel<true><<<blocks, 256, 0, s>>>(nf, th->nfp, th->d_nz, th->d_grid, th->d_s2,
                                             im->d_hdr, im->d_coef, th->d_goff, th->emiss, Fgeo,
                                             nsteps, dQ, th->d_Qprev, th->d_T, th->d_F,
                                             th->d_work, th->d_work32, Tsurf_series,
                                             th->d_status);
    else
      cn_kernel<false><<<blocks, 256, 0, s>>>(nf, th->nfp, th->d_nz, th->d_grid, th->d_s2,
                                              im->d_hdr, im->d_coef, th->d_goff, th->emiss, Fgeo,
                                              nsteps, dQ, th->d_Qprev, th->d_T, th->d_F,
                                              th->d_work, th->d_work32, Tsurf_series,
                                              th->d_status);
    ILLUM_LAUNCH_CHECK();
  }
  copy_qprev_kernel<<<blocks, 256, 0, s>>>(nf, dQ + (int64_t)(nsteps - 1) * nf, th->d_Qprev);
  ILLUM_LAUNCH_CHECK();
  int st = 0;
  ILLUM_HIP(hipMemcpyAsync(&st, th->d_status, 4, hipMemcpyDeviceToHost, s));
  sync(ctx);
  if (st) throw Error(ILLUM_ERR_ZERO_PIVOT, "tridag: rewrite equations (b(1) == 0)");
}

void thermal_get(illum_thermal *th, double *T, double *Tsurf, double *Fsurf, int flags) {
  illum_ctx *ctx = th->ctx;
  hipStream_t s = stream_of(ctx);
  const int64_t nf = th->nf;
  const bool dev = flags & ILLUM_OUT_ON_DEVICE;
  const auto kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (T) {
    const int64_t tot = nf * (th->nzmax + 1);
    DBuf<double> tmp;
    double *dst = T;
    if (!dev) { tmp.reset(tot); dst = tmp.p; }
    gather_T_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(nf, th->nfp, th->nzmax,
                                                                th->d_nz, th->d_T, dst);
    ILLUM_LAUNCH_CHECK();
    if (!dev) ILLUM_HIP(hipMemcpyAsync(T, dst, 8 * tot, hipMemcpyDeviceToHost, s));
    sync(ctx);
  }
  if (Tsurf) ILLUM_HIP(hipMemcpyAsync(Tsurf, th->d_T, 8 * nf, kind, s));
  if (Fsurf) ILLUM_HIP(hipMemcpyAsync(Fsurf, th->d_F, 8 * nf, kind, s));
  sync(ctx);
}

}  // namespace illum
