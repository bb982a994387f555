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
//   ILLUM_RAD_MIXED: iterative refinement with a one-sweep correction.  The
//     FP64 residual r = E - (I-F)B is carried from sweep to sweep; each sweep
//     rounds it to d = fp32(r), adds d to B (FP64) and updates r <- (r - d) +
//     F d with FP64 values and accumulation, gathering d in FP32.
// Both report max_s ||E_s - (I-F)B_s||_2 / ||E_s||_2, measured in FP64.
#include "hip_util.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace illum {

// Sparse products over the CSR rows of F (values TV) with a gathered block X
// (TX) and right-hand side R / output Y (TY; accumulation in TY):
//   Y = R + F X            (subtract_x = 0)
//   Y = (R - X) + F X      (subtract_x = 1: the residual E - (I - F) X for R = E)
// Rows of X, R and Y hold S contiguous values (row-major [nf][S]).  The rows
// and wide kernels sum a row's nonzeros in CSR order (identical results); the
// small-S kernel uses a fixed lane-strided order and tree.  Rows go to XCDs in contiguous ranges: neighbouring facets
// share most of their visible set, so gathered rows of X are reused from the
// XCD's L2.

template <class TY, class TX>
__device__ __forceinline__ TY spmm_finish(TY r, TX x, TY acc, int subtract_x) {
  return subtract_x ? (r - (TY)x) + acc : r + acc;
}

// One wave per row, lane s covers one right-hand side per 64-wide chunk.
template <class TV, class TX, class TY>
__global__ __launch_bounds__(256) void spmm_rows_kernel(
    int64_t nf, int S, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
    const TV *__restrict__ val, const TX *__restrict__ X, const TY *__restrict__ R,
    TY *__restrict__ Y, int subtract_x) {
  const int64_t row = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nf) return;
  const int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
  for (int s0 = 0; s0 < S; s0 += 64) {
    const int s = s0 + lane;
    if (s >= S) break;
    TY acc = 0;
    for (int64_t k = k0; k < k1; ++k)
      acc = acc + (TY)val[k] * (TY)X[(int64_t)colidx[k] * S + s];
    Y[row * S + s] = spmm_finish(R[row * S + s], X[row * S + s], acc, subtract_x);
  }
}

// Wide form for S >= 128: lane l covers PER consecutive right-hand sides (one
// vector load of PER values per nonzero), so one pass of a wave reads 64 * PER
// values of every gathered row of X and the row's indices and values once.
// Requires S % PER == 0 (vector alignment).
template <class TV, class TX, class TY, int PER>
__global__ __launch_bounds__(256) void spmm_wide_kernel(
    int64_t nf, int S, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
    const TV *__restrict__ val, const TX *__restrict__ X, const TY *__restrict__ R,
    TY *__restrict__ Y, int subtract_x) {
  typedef TX xvec __attribute__((ext_vector_type(PER)));
  typedef TY yvec __attribute__((ext_vector_type(PER)));
  const int64_t row = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nf) return;
  const int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
  for (int c0 = 0; c0 < S; c0 += 64 * PER) {
    const int s = c0 + lane * PER;
    if (s >= S) break;
    yvec acc = (yvec)(TY)0;
#pragma unroll 2
    for (int64_t k = k0; k < k1; ++k) {
      const TY f = (TY)val[k];
      const xvec x = *(const xvec *)(X + (int64_t)colidx[k] * S + s);
      acc = acc + f * __builtin_convertvector(x, yvec);
    }
    const yvec r = *(const yvec *)(R + row * S + s);
    yvec y;
    if (subtract_x)
      y = (r - __builtin_convertvector(*(const xvec *)(X + row * S + s), yvec)) + acc;
    else
      y = r + acc;
    *(yvec *)(Y + row * S + s) = y;
  }
}

// Small S (<= 31): one wave per row, lanes over nonzeros, fixed-shape wave
// reduction per right-hand side.
template <class TV, class TX, class TY>
__global__ __launch_bounds__(256) void spmm_nnz_kernel(
    int64_t nf, int S, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
    const TV *__restrict__ val, const TX *__restrict__ X, const TY *__restrict__ R,
    TY *__restrict__ Y, int subtract_x) {
  const int64_t row = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nf) return;
  const int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
  for (int s = 0; s < S; ++s) {
    TY acc = 0;
    for (int64_t k = k0 + lane; k < k1; k += 64)
      acc = acc + (TY)val[k] * (TY)X[(int64_t)colidx[k] * S + s];
    for (int o = 32; o > 0; o >>= 1) acc = acc + __shfl_xor(acc, o);
    if (lane == 0) Y[row * S + s] = spmm_finish(R[row * S + s], X[row * S + s], acc, subtract_x);
  }
}

// Partial column sums of (A - B)^2 (B nullable) over row chunks, one chunk per
// block.  With S < 256 the block's threads split into L = 256 / S groups of S
// (thread t: column t % S, row offset t / S) and the groups are summed in a
// fixed order, so the result is deterministic.
template <class TA>
__global__ __launch_bounds__(256) void colnorm_partial_kernel(int64_t nf, int S,
                                                              const TA *__restrict__ A,
                                                              const TA *__restrict__ B,
                                                              double *__restrict__ part) {
  __shared__ double red[256];
  const int64_t G = gridDim.x;
  const int64_t R = (nf + G - 1) / G;
  const int64_t r0 = blockIdx.x * R, r1 = std::min<int64_t>(nf, r0 + R);
  const int t = threadIdx.x;
  if (S >= 256) {
    for (int s = t; s < S; s += 256) {
      double acc = 0;
      for (int64_t i = r0; i < r1; ++i) {
        double d = (double)A[i * S + s];
        if (B) d = d - (double)B[i * S + s];
        acc += d * d;
      }
      part[(int64_t)blockIdx.x * S + s] = acc;
    }
    return;
  }
  const int L = 256 / S;
  const int s = t % S, q = t / S;
  double acc = 0;
  if (q < L)
    for (int64_t i = r0 + q; i < r1; i += L) {
      double d = (double)A[i * S + s];
      if (B) d = d - (double)B[i * S + s];
      acc += d * d;
    }
  red[t] = acc;
  __syncthreads();
  if (t < S) {
    double tot = 0;
    for (int k = 0; k < L; ++k) tot += red[k * S + t];
    part[(int64_t)blockIdx.x * S + t] = tot;
  }
}

// out[s] = sum over the G partials of column s: one block per column, strided
// per-thread sums then a fixed-shape tree (deterministic).
__global__ __launch_bounds__(256) void colnorm_final_kernel(int G, int S,
                                                            const double *__restrict__ part,
                                                            double *__restrict__ out) {
  __shared__ double red[256];
  const int s = blockIdx.x;
  double acc = 0;
  for (int g = threadIdx.x; g < G; g += 256) acc += part[(int64_t)g * S + s];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[s] = red[0];
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

// Iterative-refinement update: d = (float) r, B += d.
__global__ void refine_kernel(int64_t n, const double *__restrict__ r, float *__restrict__ d,
                              double *__restrict__ B) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float dk = (float)r[k];
  d[k] = dk;
  B[k] = B[k] + (double)dk;
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
  template <class TV, class TX, class TY>
  void spmm(const illum_sparse *F, const TV *val, const TX *X, const TY *R, TY *Y, int subx,
            const char *stage) {
    StageTimer st(ctx, stage);
    const unsigned blocks = (unsigned)((nf + 3) / 4);
#define ILLUM_SPMM_WIDE(P)                                                                  \
  spmm_wide_kernel<TV, TX, TY, P><<<blocks, 256, 0, s>>>(nf, S, F->d_rowptr, F->d_colidx, val, X, \
                                                     R, Y, subx)
    if (S >= 512 && S % 8 == 0)
      ILLUM_SPMM_WIDE(8);
    else if (S >= 256 && S % 4 == 0)
      ILLUM_SPMM_WIDE(4);
    else if (S >= 128 && S % 2 == 0)
      ILLUM_SPMM_WIDE(2);
#undef ILLUM_SPMM_WIDE
    else if (S >= 32)
      spmm_rows_kernel<TV, TX, TY><<<blocks, 256, 0, s>>>(nf, S, F->d_rowptr, F->d_colidx, val, X,
                                                      R, Y, subx);
    else
      spmm_nnz_kernel<TV, TX, TY><<<blocks, 256, 0, s>>>(nf, S, F->d_rowptr, F->d_colidx, val, X,
                                                     R, Y, subx);
    ILLUM_LAUNCH_CHECK();
  }
  template <class TA>
  std::vector<double> colnorm2(const TA *A, const TA *B) {
    colnorm_partial_kernel<TA><<<G, 256, 0, s>>>(nf, S, A, B, part.p);
    ILLUM_LAUNCH_CHECK();
    colnorm_final_kernel<<<S, 256, 0, s>>>(G, S, part.p, out.p);
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
      c.spmm<double, double, double>(F, F->d_rval, cur, Er.p, nxt, 0, "radiosity_fp64_sweep");
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
    // Iterative refinement with a one-sweep FP32 correction: d = fp32(r),
    // B += d, r <- (r - d) + F d with FP64 values and FP64 accumulation.  r
    // stays the FP64 residual E - (I - F) B (r - d is exact: d is r rounded),
    // and |r_{k+1}| <= (rho(F) + 2^-24) |r_k|: the FP64 Jacobi rate, while the
    // gathered block (S values per nonzero) moves in FP32.
    DBuf<double> R0(n), R1(n);
    DBuf<float> D(n);
    ILLUM_HIP(hipMemsetAsync(B0.p, 0, 8 * n, s));
    ILLUM_HIP(hipMemcpyAsync(R0.p, Er.p, 8 * n, hipMemcpyDeviceToDevice, s));
    double *r = R0.p, *rn = R1.p;
    const unsigned eb = (unsigned)((n + 255) / 256);
    while (nprod < max_iter) {
      refine_kernel<<<eb, 256, 0, s>>>(n, r, D.p, B0.p);
      ILLUM_LAUNCH_CHECK();
      c.spmm<double, float, double>(F, F->d_rval, D.p, r, rn, 1, "radiosity_mixed_sweep");
      ++nprod;
      rel = max_rel(c.colnorm2<double>(rn, nullptr), e2);
      record(rel);
      std::swap(r, rn);
      if (rel <= rel_tol) break;
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
