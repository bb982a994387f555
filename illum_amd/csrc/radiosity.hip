// Radiosity solve (I - F) B = E for many sun directions at once.
//
// Reference: the Gauss-Seidel loop of do_radiosity_task (illum_cli.cpp:425-440,
// arma_util.cpp:3-19) -- serial, one right-hand side.  Gauss-Seidel has a
// sequential dependency along facets; on the GPU this library uses Jacobi
// sweeps B <- E + F B, which share Gauss-Seidel's fixed point and converge
// because rho(F) <= max albedo < 1, and solves all S right-hand sides in one
// sparse-times-dense product per sweep (S values of one facet are contiguous,
// so every nonzero F(i,j) gathers one contiguous row of B).
//
//   ILLUM_RAD_JACOBI_FP64: FP64 values, FP64 vectors.
//   ILLUM_RAD_MIXED: iterative refinement.  The residual r = E - (I-F)B is
//     formed in FP64 (FP64 values, FP64 B); the correction (I-F) d = r is
//     solved by FP32 Jacobi sweeps on FP32 values; B += d in FP64.
// Both report max_s ||E_s - (I-F)B_s||_2 / ||E_s||_2, measured in FP64.
#include "hip_util.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace illum {

// Y = R + F X - c X  (rows of X, R, Y hold S contiguous values), one wave per
// row, lanes over right-hand sides.  Used when S >= 32.
template <class TV, class TX>
__global__ __launch_bounds__(256) void spmm_rows_kernel(
    int64_t nf, int S, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
    const TV *__restrict__ val, const TX *__restrict__ X, const TX *__restrict__ R,
    TX *__restrict__ Y, int subtract_x) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nf) return;
  const int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
  for (int s0 = 0; s0 < S; s0 += 64) {
    const int s = s0 + lane;
    TX acc = 0;
    if (s < S) {
      for (int64_t k = k0; k < k1; ++k) {
        const TX f = (TX)val[k];
        acc = acc + f * X[(int64_t)colidx[k] * S + s];
      }
      TX y = R[row * S + s] + acc;
      if (subtract_x) y = y - X[row * S + s];
      Y[row * S + s] = y;
    }
  }
}

// Same product for small S: one wave per row, lanes over nonzeros, wave
// reduction per right-hand side (S <= 31).
template <class TV, class TX>
__global__ __launch_bounds__(256) void spmm_nnz_kernel(
    int64_t nf, int S, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
    const TV *__restrict__ val, const TX *__restrict__ X, const TX *__restrict__ R,
    TX *__restrict__ Y, int subtract_x) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nf) return;
  const int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
  for (int s = 0; s < S; ++s) {
    TX acc = 0;
    for (int64_t k = k0 + lane; k < k1; k += 64) acc = acc + (TX)val[k] * X[(int64_t)colidx[k] * S + s];
    for (int o = 32; o > 0; o >>= 1) acc = acc + __shfl_xor(acc, o);
    if (lane == 0) {
      TX y = R[row * S + s] + acc;
      if (subtract_x) y = y - X[row * S + s];
      Y[row * S + s] = y;
    }
  }
}

// Partial column sums of (A - B)^2 (B nullable) over row chunks; deterministic.
template <class TA>
__global__ void colnorm_partial_kernel(int64_t nf, int S, const TA *__restrict__ A,
                                       const TA *__restrict__ B, double *__restrict__ part) {
  const int64_t G = gridDim.x;
  const int64_t R = (nf + G - 1) / G;
  const int64_t r0 = blockIdx.x * R, r1 = std::min<int64_t>(nf, r0 + R);
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    double acc = 0;
    for (int64_t i = r0; i < r1; ++i) {
      double d = (double)A[i * S + s];
      if (B) d = d - (double)B[i * S + s];
      acc += d * d;
    }
    part[(int64_t)blockIdx.x * S + s] = acc;
  }
}

__global__ void colnorm_final_kernel(int G, int S, const double *__restrict__ part,
                                     double *__restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  double acc = 0;
  for (int g = 0; g < G; ++g) acc += part[(int64_t)g * S + s];
  out[s] = acc;
}

// Column-major (nf x S) <-> row-major ([nf][S]) copies with optional cast.
template <class TI, class TO>
__global__ void transpose_kernel(int64_t nf, int S, const TI *__restrict__ in,
                                 TO *__restrict__ out, int to_rowmajor) {
  __shared__ TI tile[32][33];
  const int64_t i0 = (int64_t)blockIdx.x * 32;
  const int s0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    if (to_rowmajor) {  // in: col-major [s][i]
      const int64_t i = i0 + tx;
      const int s = s0 + k;
      if (i < nf && s < S) tile[k][tx] = in[(int64_t)s * nf + i];
    } else {            // in: row-major [i][s]
      const int64_t i = i0 + k;
      const int s = s0 + tx;
      if (i < nf && s < S) tile[k][tx] = in[i * S + s];
    }
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    if (to_rowmajor) {  // out[i][s]: i = i0 + k, s = s0 + tx  (tile[s][i])
      const int64_t i = i0 + k;
      const int s = s0 + tx;
      if (i < nf && s < S) out[i * S + s] = (TO)tile[tx][k];
    } else {            // out[s][i]: s = s0 + k, i = i0 + tx  (tile[i][s])
      const int64_t i = i0 + tx;
      const int s = s0 + k;
      if (i < nf && s < S) out[(int64_t)s * nf + i] = (TO)tile[tx][k];
    }
  }
}

template <class TI, class TO>
__global__ void axpy_cast_kernel(int64_t n, const TI *__restrict__ x, TO *__restrict__ y,
                                 int accumulate) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (accumulate) y[k] = y[k] + (TO)x[k];
  else y[k] = (TO)x[k];
}

namespace {

struct Ctx {
  illum_ctx *ctx;
  hipStream_t s;
  int64_t nf;
  int S;
  int G;
  DBuf<double> part, out;
  Ctx(illum_ctx *c, int64_t n, int S_) : ctx(c), s(stream_of(c)), nf(n), S(S_) {
    G = (int)std::min<int64_t>(1024, std::max<int64_t>(1, n / 16));
    part.reset((size_t)G * S);
    out.reset(S);
  }
  template <class TV, class TX>
  void spmm(const illum_sparse *F, const TV *val, const TX *X, const TX *R, TX *Y, int subx,
            const char *stage) {
    StageTimer st(ctx, stage);
    const unsigned blocks = (unsigned)((nf + 3) / 4);
    if (S >= 32)
      spmm_rows_kernel<TV, TX><<<blocks, 256, 0, s>>>(nf, S, F->d_rowptr, F->d_colidx, val, X,
                                                      R, Y, subx);
    else
      spmm_nnz_kernel<TV, TX><<<blocks, 256, 0, s>>>(nf, S, F->d_rowptr, F->d_colidx, val, X,
                                                     R, Y, subx);
    ILLUM_LAUNCH_CHECK();
  }
  template <class TA>
  std::vector<double> colnorm2(const TA *A, const TA *B) {
    colnorm_partial_kernel<TA><<<G, 256, 0, s>>>(nf, S, A, B, part.p);
    ILLUM_LAUNCH_CHECK();
    colnorm_final_kernel<<<(S + 255) / 256, 256, 0, s>>>(G, S, part.p, out.p);
    ILLUM_LAUNCH_CHECK();
    std::vector<double> h(S);
    ILLUM_HIP(hipMemcpyAsync(h.data(), out.p, 8 * S, hipMemcpyDeviceToHost, s));
    sync(ctx);
    return h;
  }
};

double max_rel(const std::vector<double> &r2, const std::vector<double> &e2) {
  double m = 0;
  for (size_t s = 0; s < r2.size(); ++s) {
    const double rel = e2[s] > 0 ? std::sqrt(r2[s] / e2[s]) : std::sqrt(r2[s]);
    m = std::max(m, rel);
  }
  return m;
}

template <class TI, class TO>
void transpose(hipStream_t s, int64_t nf, int S, const TI *in, TO *out, bool to_rm) {
  dim3 g((unsigned)((nf + 31) / 32), (unsigned)((S + 31) / 32));
  transpose_kernel<TI, TO><<<g, 256, 0, s>>>(nf, S, in, out, to_rm ? 1 : 0);
  ILLUM_LAUNCH_CHECK();
}

}  // namespace

void radiosity_solve(illum_sparse *F, const double *E, int S, int method, double rel_tol,
                     int max_iter, double *B, double *resid, int *iters, int flags) {
  illum_ctx *ctx = F->ctx;
  if (!F->d_rowptr) throw Error(ILLUM_ERR_ARG, "matrix has no CSR form");
  if (S < 1 || max_iter < 1 || !(rel_tol >= 0)) throw Error(ILLUM_ERR_ARG, "bad solve parameters");
  if (method != ILLUM_RAD_JACOBI_FP64 && method != ILLUM_RAD_MIXED)
    throw Error(ILLUM_ERR_ARG, "unknown radiosity method");
  if (!E || !B) throw Error(ILLUM_ERR_ARG, "null pointer");
  const int64_t nf = F->n;
  const int64_t n = nf * S;
  Ctx c(ctx, nf, S);
  hipStream_t s = c.s;

  // Inputs to row-major FP64.
  DBuf<double> Ecm;
  const double *Ein = E;
  if (!(flags & ILLUM_IN_ON_DEVICE)) {
    Ecm.reset(n);
    ILLUM_HIP(hipMemcpyAsync(Ecm.p, E, 8 * n, hipMemcpyHostToDevice, s));
    Ein = Ecm.p;
  }
  DBuf<double> Er(n), B0(n), B1(n);
  if (S == 1) ILLUM_HIP(hipMemcpyAsync(Er.p, Ein, 8 * n, hipMemcpyDeviceToDevice, s));
  else transpose(s, nf, S, Ein, Er.p, true);
  const std::vector<double> e2 = c.colnorm2<double>(Er.p, nullptr);

  int nprod = 0, nhist = 0;
  double rel = 1.0;
  double *Bcur = nullptr;
  auto record = [&](double r) {
    if (resid && nhist < max_iter + 1) resid[nhist] = r;
    ++nhist;
  };

  if (method == ILLUM_RAD_JACOBI_FP64) {
    // B_1 = E (the sweep from B_0 = 0); then B_{k+1} = E + F B_k.
    ILLUM_HIP(hipMemcpyAsync(B0.p, Er.p, 8 * n, hipMemcpyDeviceToDevice, s));
    double *cur = B0.p, *nxt = B1.p;
    bool ok = false;
    while (nprod < max_iter) {
      c.spmm<double, double>(F, F->d_rval, cur, Er.p, nxt, 0, "radiosity_fp64_sweep");
      ++nprod;
      rel = max_rel(c.colnorm2<double>(nxt, cur), e2);  // residual of cur
      record(rel);
      if (rel <= rel_tol) { ok = true; break; }
      std::swap(cur, nxt);
    }
    Bcur = cur;
    if (!ok) {
      // cur now holds the newest iterate; report its residual bound honestly
      Bcur = cur;
    }
    if (iters) *iters = nprod;
  } else {
    // Iterative refinement with FP32 inner Jacobi sweeps.
    DBuf<double> Rr(n);
    DBuf<float> R32(n), D0(n), D1(n);
    ILLUM_HIP(hipMemsetAsync(B0.p, 0, 8 * n, s));
    bool first = true;
    int nouter = 0;
    while (true) {
      if (first) {
        ILLUM_HIP(hipMemcpyAsync(Rr.p, Er.p, 8 * n, hipMemcpyDeviceToDevice, s));
        rel = 1.0;
        first = false;
      } else {
        c.spmm<double, double>(F, F->d_rval, B0.p, Er.p, Rr.p, 1, "radiosity_fp64_residual");
        ++nprod;
        rel = max_rel(c.colnorm2<double>(Rr.p, nullptr), e2);
      }
      record(rel);
      if (rel <= rel_tol || nprod >= max_iter) break;
      ++nouter;
      // d = r + F d, FP32, until the update is below 1e-6 of ||r||.
      const unsigned eb = (unsigned)((n + 255) / 256);
      axpy_cast_kernel<double, float><<<eb, 256, 0, s>>>(n, Rr.p, R32.p, 0);
      ILLUM_LAUNCH_CHECK();
      ILLUM_HIP(hipMemcpyAsync(D0.p, R32.p, 4 * n, hipMemcpyDeviceToDevice, s));
      const std::vector<double> r2 = c.colnorm2<float>(R32.p, nullptr);
      float *dc = D0.p, *dn = D1.p;
      for (int inner = 0; inner < 64 && nprod < max_iter; ++inner) {
        c.spmm<float, float>(F, F->d_rval32, dc, R32.p, dn, 0, "radiosity_fp32_sweep");
        ++nprod;
        const double ch = max_rel(c.colnorm2<float>(dn, dc), r2);
        std::swap(dc, dn);
        if (ch <= 1e-6) break;
      }
      axpy_cast_kernel<float, double><<<eb, 256, 0, s>>>(n, dc, B0.p, 1);
      ILLUM_LAUNCH_CHECK();
    }
    Bcur = B0.p;
    if (iters) *iters = nprod;
  }

  // Output to column-major.
  DBuf<double> Bout;
  double *dst = B;
  if (!(flags & ILLUM_OUT_ON_DEVICE)) { Bout.reset(n); dst = Bout.p; }
  if (S == 1) ILLUM_HIP(hipMemcpyAsync(dst, Bcur, 8 * n, hipMemcpyDeviceToDevice, s));
  else transpose(s, nf, S, Bcur, dst, false);
  if (!(flags & ILLUM_OUT_ON_DEVICE))
    ILLUM_HIP(hipMemcpyAsync(B, dst, 8 * n, hipMemcpyDeviceToHost, s));
  sync(ctx);
  if (rel > rel_tol)
    throw Error(ILLUM_ERR_NOT_CONVERGED,
                "radiosity solve did not reach the tolerance (relative residual " +
                    std::to_string(rel) + ")");
}

}  // namespace illum
