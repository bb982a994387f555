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
#include <cstdlib>
#include <cstring>
#include <vector>

namespace illum {

// Sparse products over the CSR rows of F (values TV) with a gathered block X
// (TX) and right-hand side R / output Y (FP64 accumulation):
//   Y = R + F X            (subtract_x = 0)
//   Y = (R - X) + F X      (subtract_x = 1: the residual E - (I - F) X for R = E)
// Rows of X, R and Y hold S contiguous values (row-major [nf][S]).  The rows
// and wide kernels sum a row's nonzeros in CSR order (identical results); the
// small-S kernel uses a fixed lane-strided order and tree.  Rows go to XCDs in
// contiguous ranges: neighbouring facets share most of their visible set, so
// gathered rows of X are reused from the XCD's L2.
//
// Each product also does the per-element work of the sweep around it, on the
// rows it owns (no other launch touches the vectors in between):
//   B  != null: B += X at the row (the mixed method's B += d);
//   Dn != null: Dn = fp32(Y)      (the mixed method's next correction d);
//   part != null (small-S kernel): per-block column sums of (Y - X)^2
//               (norm_sub_x) or Y^2, for the residual norm.
struct SweepOut {
  double *B;
  float *Dn;
  double *part;
  int norm_sub_x;
};

__device__ __forceinline__ void sweep_elem(const SweepOut &o, int64_t at, double x, double y) {
  if (o.B) o.B[at] = o.B[at] + x;
  if (o.Dn) o.Dn[at] = (float)y;
}

// One wave per row, lane s covers one right-hand side per 64-wide chunk.
template <class TV, class TX>
__global__ __launch_bounds__(256) void spmm_rows_kernel(
    int64_t nf, int S, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
    const TV *__restrict__ val, const TX *__restrict__ X, const double *__restrict__ R,
    double *__restrict__ Y, int subtract_x, const int *__restrict__ done, SweepOut o) {
  if (*done) return;
  const int64_t row = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nf) return;
  const int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
  for (int s0 = 0; s0 < S; s0 += 64) {
    const int s = s0 + lane;
    if (s >= S) break;
    double acc = 0;
    for (int64_t k = k0; k < k1; ++k)
      acc = acc + (double)val[k] * (double)X[(int64_t)colidx[k] * S + s];
    const int64_t at = row * S + s;
    const double x = (double)X[at];
    const double y = subtract_x ? (R[at] - x) + acc : R[at] + acc;
    Y[at] = y;
    sweep_elem(o, at, x, y);
  }
}

// Wide form: lane l covers PER consecutive right-hand sides (one vector load
// of PER values per nonzero).  The grid is column-chunk-major: chunk ch
// (columns [64 PER ch, 64 PER (ch + 1))) runs over all rows before the next
// chunk starts, so the gathered block of X in use is nf x 64 PER values,
// sized by the host to stay resident in the memory-side cache; per chunk
// nrb8 blocks (a multiple of 8, so each XCD's blocks take a contiguous row
// range).  Requires S % PER == 0 (vector alignment).  Each output element
// sums its row's nonzeros in CSR order, whatever PER.
template <class TV, class TX, int PER>
__global__ __launch_bounds__(256) void spmm_wide_kernel(
    int64_t nf, int S, int nrb8, const int64_t *__restrict__ rowptr,
    const int32_t *__restrict__ colidx, const TV *__restrict__ val, const TX *__restrict__ X,
    const double *__restrict__ R, double *__restrict__ Y, int subtract_x,
    const int *__restrict__ done, SweepOut o) {
  if (*done) return;
  typedef TX xvec __attribute__((ext_vector_type(PER)));
  typedef double yvec __attribute__((ext_vector_type(PER)));
  typedef float fvec __attribute__((ext_vector_type(PER)));
  const int ch = blockIdx.x / nrb8, rb = blockIdx.x - ch * nrb8;
  const int64_t row = (int64_t)xcd_remap(rb, nrb8) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int s = (ch * 64 + lane) * PER;
  if (row >= nf || s >= S) return;
  const int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
  yvec acc = (yvec)0.0;
#pragma unroll 4
  for (int64_t k = k0; k < k1; ++k) {
    const double f = (double)val[k];
    const xvec x = *(const xvec *)(X + (int64_t)colidx[k] * S + s);
    acc = acc + f * __builtin_convertvector(x, yvec);
  }
  const int64_t at = row * S + s;
  const yvec r = *(const yvec *)(R + at);
  const yvec x = __builtin_convertvector(*(const xvec *)(X + at), yvec);
  const yvec y = subtract_x ? (r - x) + acc : r + acc;
  *(yvec *)(Y + at) = y;
  if (o.B) *(yvec *)(o.B + at) = *(const yvec *)(o.B + at) + x;
  if (o.Dn) *(fvec *)(o.Dn + at) = __builtin_convertvector(y, fvec);
}

// Small S (< 32): LPR lanes per row (64 / LPR rows per wave), lanes strided
// over the row's nonzeros, fixed-shape reduction per right-hand side; the
// block's column sums of squares go to o.part[block][S] (rows in order).
constexpr int SPMM_NNZ_LPR = 16;
constexpr int SPMM_NNZ_ROWS = 4 * 64 / SPMM_NNZ_LPR;  // rows per 256-thread block
constexpr int SPMM_NNZ_MAXS = 31;

template <class TV, class TX>
__global__ __launch_bounds__(256) void spmm_nnz_kernel(
    int64_t nf, int S, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
    const TV *__restrict__ val, const TX *__restrict__ X, const double *__restrict__ R,
    double *__restrict__ Y, int subtract_x, const int *__restrict__ done, SweepOut o) {
  constexpr int LPR = SPMM_NNZ_LPR;
  __shared__ double sq[SPMM_NNZ_ROWS][SPMM_NNZ_MAXS + 1];
  if (*done) return;  // uniform: the whole grid returns or none of it
  const int rib = threadIdx.x / LPR;  // row in block
  const int sub = threadIdx.x % LPR;
  const int64_t row = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * SPMM_NNZ_ROWS + rib;
  const bool valid = row < nf;
  const int64_t k0 = valid ? rowptr[row] : 0, k1 = valid ? rowptr[row + 1] : 0;
  for (int s = 0; s < S; ++s) {
    // the row's own operands are loaded ahead of the gather
    const int64_t at = row * S + s;
    const bool own = valid && sub == 0;
    const double r_at = own ? R[at] : 0.0, x_at = own ? (double)X[at] : 0.0;
    const double b_at = own && o.B ? o.B[at] : 0.0;
    double acc = 0;
#pragma unroll 4
    for (int64_t k = k0 + sub; k < k1; k += LPR)
      acc = acc + (double)val[k] * (double)X[(int64_t)colidx[k] * S + s];
    for (int w = LPR / 2; w > 0; w >>= 1) acc = acc + __shfl_xor(acc, w);
    if (sub == 0) {
      double d2 = 0.0;
      if (valid) {
        const double y = subtract_x ? (r_at - x_at) + acc : r_at + acc;
        Y[at] = y;
        if (o.B) o.B[at] = b_at + x_at;
        if (o.Dn) o.Dn[at] = (float)y;
        const double d = o.norm_sub_x ? y - x_at : y;
        d2 = d * d;
      }
      sq[rib][s] = d2;
    }
  }
  if (!o.part) return;
  __syncthreads();
  if ((int)threadIdx.x < S) {
    double t = 0.0;
    for (int r = 0; r < SPMM_NNZ_ROWS; ++r) t += sq[r][threadIdx.x];
    o.part[(int64_t)blockIdx.x * S + threadIdx.x] = t;
  }
}

// Solver state in device memory.  The stopping decision is taken on the
// device, so the host launches several sweeps ahead and waits only once per
// batch; launches after the decision return at once.  (A last-block-done
// reduction inside one launch would need an agent-scope fence per block,
// which writes back and invalidates the XCD's L2: far slower than a second
// small launch.)
struct SolveCtl {
  int done;   // the tolerance was met; later sweeps of the solve do nothing
  int nprod;  // products taken
};

// Column sums over rows [r0, r1) of a row-major [rows][S] array, for the
// columns [c0, c0 + cw), cw <= 256: of (A - B)^2 (B nullable) when SQ, else of
// A.  The block's threads split into L = 256 / cw groups of cw (thread t:
// column c0 + t % cw, rows r0 + t / cw, stepping by L), combined in a fixed
// order: deterministic.  Column s's total is written to out[s].
template <bool SQ, class TA>
__device__ void block_colsum(int64_t r0, int64_t r1, int S, int c0, int cw, const TA *A,
                             const TA *B, double *red, double *out) {
  const int t = threadIdx.x;
  const int L = 256 / cw;
  const int c = t % cw, q = t / cw;
  double acc = 0;
  if (q < L) {
    const int64_t s = c0 + c;
#pragma unroll 4
    for (int64_t i = r0 + q; i < r1; i += L) {
      double d = (double)A[i * S + s];
      if (SQ && B) d = d - (double)B[i * S + s];
      acc += SQ ? d * d : d;
    }
  }
  __syncthreads();  // red is free (a previous window's combine has finished)
  red[t] = acc;
  __syncthreads();
  if (L <= 16) {
    if (t < cw) {
      double tot = 0;
      for (int k = 0; k < L; ++k) tot += red[k * cw + t];
      out[c0 + t] = tot;
    }
    return;
  }
  // Many groups (cw < 16): 16 sub-sums of consecutive groups, then their sum.
  __shared__ double red2[256];
  const int per = (L + 15) / 16;
  if (t < 16 * cw) {
    const int j = t / cw;
    double u = 0;
    for (int k = j * per; k < min(L, (j + 1) * per); ++k) u += red[k * cw + c];
    red2[t] = u;
  }
  __syncthreads();
  if (t < cw) {
    double tot = 0;
    for (int j = 0; j < 16; ++j) tot += red2[j * cw + t];
    out[c0 + t] = tot;
  }
}

__device__ __forceinline__ double max_nan(double a, double b) {
  return (b > a || b != b) ? b : a;  // a NaN, once seen, is kept
}

// Partial column sums of (A - B)^2 (B nullable) of a row-major [nf][S]
// array, one row chunk per block, to part[block][S].  Does nothing once done
// is set (done nullable).
template <class TA>
__global__ __launch_bounds__(256) void colnorm_partial_kernel(int64_t nf, int S,
                                                              const TA *__restrict__ A,
                                                              const TA *__restrict__ B,
                                                              double *__restrict__ part,
                                                              const int *__restrict__ done) {
  if (done && *done) return;
  __shared__ double red[256];
  const int64_t G = gridDim.x;
  const int64_t R = (nf + G - 1) / G;
  const int64_t r0 = std::min<int64_t>(nf, blockIdx.x * R), r1 = std::min<int64_t>(nf, r0 + R);
  for (int c0 = 0; c0 < S; c0 += 256)
    block_colsum<true>(r0, r1, S, c0, std::min(256, S - c0), A, B, red,
                       part + (int64_t)blockIdx.x * S);
}

// The stop decision: rel = max_s sqrt(col_s / e2_s) (sqrt(col_s) where e2_s =
// 0) is stored as hist[nprod], the product counted, done set when rel <= tol.
// Called by one whole block after col[] is complete.
__device__ void decide(int S, const double *col, const double *e2, SolveCtl *ctl, double *hist,
                       double tol, double *red) {
  double m = 0.0;
  for (int s = threadIdx.x; s < S; s += 256) {
    const double r2 = col[s];
    m = max_nan(m, e2[s] > 0 ? sqrt(r2 / e2[s]) : sqrt(r2));
  }
  __syncthreads();
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = max_nan(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double rel = red[0];
    const int k = ctl->nprod;
    hist[k] = rel;
    ctl->nprod = k + 1;
    if (rel <= tol) ctl->done = 1;
  }
}

// Adds the G partials of columns [b * COLNORM_CW, +COLNORM_CW) (block b, all
// columns when S <= COLNORM_CW) in a fixed order into col[].  record = 1:
// nothing once done is set, and a single block also takes the stop decision
// (with several blocks decide_kernel follows).
constexpr int COLNORM_CW = 32;

__global__ __launch_bounds__(256) void colnorm_final_kernel(int G, int S,
                                                            const double *__restrict__ part,
                                                            double *__restrict__ col,
                                                            const double *__restrict__ e2,
                                                            SolveCtl *__restrict__ ctl,
                                                            double *__restrict__ hist, double tol,
                                                            int record) {
  if (record && ctl->done) return;
  __shared__ double red[256];
  const int c0 = blockIdx.x * COLNORM_CW;
  block_colsum<false>(0, G, S, c0, std::min(COLNORM_CW, S - c0), part, (const double *)nullptr,
                      red, col);
  if (record && gridDim.x == 1) {
    __syncthreads();
    decide(S, col, e2, ctl, hist, tol, red);
  }
}

__global__ __launch_bounds__(256) void decide_kernel(int S, const double *__restrict__ col,
                                                     const double *__restrict__ e2,
                                                     SolveCtl *__restrict__ ctl,
                                                     double *__restrict__ hist, double tol) {
  if (ctl->done) return;
  __shared__ double red[256];
  decide(S, col, e2, ctl, hist, tol, red);
}

__global__ void solve_init_kernel(SolveCtl *ctl) {
  ctl->done = 0;
  ctl->nprod = 0;
}

// Mixed method start: B = 0, d_0 = fp32(E).
__global__ void mixed_init_kernel(int64_t n, const double *__restrict__ E,
                                  float *__restrict__ D, double *__restrict__ B) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  D[k] = (float)E[k];
  B[k] = 0.0;
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

namespace {

// Sparse product dispatch by S (see the kernels above).  Returns the number
// of column-sum partials written to o.part (0: none, the small-S kernel only).
template <class TV, class TX>
int spmm(hipStream_t s, illum_ctx *ctx, const illum_sparse *F, int S, const TV *val, const TX *X,
         const double *R, double *Y, int subx, const int *done, SweepOut o, const char *stage) {
  StageTimer st(ctx, stage);
  const int64_t nf = F->n;
  const unsigned blocks = (unsigned)((nf + 3) / 4);
  int nparts = 0;
  if (S >= 128) {
    // PER: the widest of 8, 4, 2, 1 dividing S whose gathered chunk
    // (nf x 64 PER values of X) stays within 128 MB, so it is served from the
    // 256 MB memory-side cache rather than HBM.
    // ILLUM_SPMM_CHUNK_BYTES overrides the 128 MB budget.
    const char *env = getenv("ILLUM_SPMM_CHUNK_BYTES");
    const double budget = env ? atof(env) : 128e6;
    int per = 8;
    while (per > 1 && (S % per != 0 || (double)nf * 64 * per * sizeof(TX) > budget)) per >>= 1;
    const int nrb8 = (int)(((nf + 3) / 4 + 7) / 8 * 8);
    const int nch = (S + 64 * per - 1) / (64 * per);
    const unsigned grid = (unsigned)((int64_t)nrb8 * nch);
#define ILLUM_SPMM_WIDE(P)                                                                   \
  spmm_wide_kernel<TV, TX, P><<<grid, 256, 0, s>>>(nf, S, nrb8, F->d_rowptr, F->d_colidx, val, X, \
                                                   R, Y, subx, done, o)
    if (per == 8) ILLUM_SPMM_WIDE(8);
    else if (per == 4) ILLUM_SPMM_WIDE(4);
    else if (per == 2) ILLUM_SPMM_WIDE(2);
    else ILLUM_SPMM_WIDE(1);
#undef ILLUM_SPMM_WIDE
  } else if (S > SPMM_NNZ_MAXS) {
    spmm_rows_kernel<TV, TX><<<blocks, 256, 0, s>>>(nf, S, F->d_rowptr, F->d_colidx, val, X, R, Y,
                                                    subx, done, o);
  } else {
    const unsigned nb = (unsigned)((nf + SPMM_NNZ_ROWS - 1) / SPMM_NNZ_ROWS);
    spmm_nnz_kernel<TV, TX><<<nb, 256, 0, s>>>(nf, S, F->d_rowptr, F->d_colidx, val, X, R, Y, subx,
                                               done, o);
    nparts = o.part ? (int)nb : 0;
  }
  ILLUM_LAUNCH_CHECK();
  return nparts;
}

template <class TI, class TO>
void transpose(hipStream_t s, int64_t nf, int S, const TI *in, TO *out, bool to_rm) {
  dim3 g((unsigned)((nf + 31) / 32), (unsigned)((S + 31) / 32));
  transpose_kernel<TI, TO><<<g, 256, 0, s>>>(nf, S, in, out, to_rm ? 1 : 0);
  ILLUM_LAUNCH_CHECK();
}

bool overlap(const void *a, const void *b, size_t bytes) {
  const char *x = (const char *)a, *y = (const char *)b;
  return x < y + bytes && y < x + bytes;
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
  hipStream_t s = stream_of(ctx);
  const bool in_dev = flags & ILLUM_IN_ON_DEVICE, out_dev = flags & ILLUM_OUT_ON_DEVICE;
  const bool mixed = method == ILLUM_RAD_MIXED;

  // Scratch from the context's workspace (no allocation in a time loop):
  // control block, residual history, column norms and partial sums.
  const int G = (int)std::min<int64_t>(1024, std::max<int64_t>(1, n / 16));
  const int64_t gparts = std::max<int64_t>(G, (nf + SPMM_NNZ_ROWS - 1) / SPMM_NNZ_ROWS);
  auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_hist = 256, o_e2 = o_hist + align(8 * ((size_t)max_iter + 1)),
               o_col = o_e2 + align(8 * (size_t)S), o_part = o_col + align(8 * (size_t)S),
               ctl_bytes = o_part + 8 * (size_t)gparts * S;
  char *ws = (char *)ctx_workspace(ctx, WS_RAD_CTL, ctl_bytes);
  SolveCtl *ctl = (SolveCtl *)ws;
  double *hist = (double *)(ws + o_hist), *e2 = (double *)(ws + o_e2),
         *col = (double *)(ws + o_col), *part = (double *)(ws + o_part);
  const int *done = &ctl->done;

  // Right-hand sides, row-major FP64.  With S = 1 both layouts coincide and a
  // device E is read in place (it is never written).
  DBuf<double> Ehost;
  const double *Ein = E;
  if (!in_dev) {
    Ehost.reset(n);
    ILLUM_HIP(hipMemcpyAsync(Ehost.p, E, 8 * n, hipMemcpyHostToDevice, s));
    Ein = Ehost.p;
  }
  const double *Er = Ein;
  if (S > 1) {
    double *t = (double *)ctx_workspace(ctx, WS_RAD_E, 8 * n);
    transpose(s, nf, S, Ein, t, true);
    Er = t;
  }
  double *X0 = (double *)ctx_workspace(ctx, WS_RAD_X0, 8 * n);
  double *X1 = (double *)ctx_workspace(ctx, WS_RAD_X1, 8 * n);
  // The mixed method accumulates B in place in the caller's B when it can
  // (device, S = 1, not overlapping E).
  double *Bacc = nullptr;
  float *D0 = nullptr, *D1 = nullptr;
  if (mixed) {
    Bacc = out_dev && S == 1 && !overlap(B, Ein, 8 * n)
               ? B : (double *)ctx_workspace(ctx, WS_RAD_X2, 8 * n);
    D0 = (float *)ctx_workspace(ctx, WS_RAD_D32, 8 * n);
    D1 = D0 + n;
  }

  solve_init_kernel<<<1, 1, 0, s>>>(ctl);
  ILLUM_LAUNCH_CHECK();
  colnorm_partial_kernel<double><<<G, 256, 0, s>>>(nf, S, Er, nullptr, part, nullptr);
  ILLUM_LAUNCH_CHECK();
  const unsigned fb = (unsigned)((S + COLNORM_CW - 1) / COLNORM_CW);
  colnorm_final_kernel<<<fb, 256, 0, s>>>(G, S, part, e2, nullptr, ctl, nullptr, 0.0, 0);
  ILLUM_LAUNCH_CHECK();
  if (mixed) {
    mixed_init_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, Er, D0, Bacc);
    ILLUM_LAUNCH_CHECK();
  }

  // Sweep k's buffers.  FP64 Jacobi, B_{k+1} = E + F B_k from B_0 = E:
  // cur_k = E, X1, X0, X1, ...; nxt_k = X1, X0, X1, ...  Mixed refinement,
  // r_0 = E: r_k = E, X0, X1, X0, ...; r_{k+1} = X0, X1, X0, ...; d_k
  // alternates D0, D1.
  auto cur = [&](int k) -> const double * { return k == 0 ? Er : (k & 1 ? X1 : X0); };
  auto nxt = [&](int k) -> double * { return k & 1 ? X0 : X1; };
  auto rk = [&](int k) -> const double * { return k == 0 ? Er : (k & 1 ? X0 : X1); };
  auto rk1 = [&](int k) -> double * { return k & 1 ? X1 : X0; };
  auto sweep = [&](int k) {
    int nparts;
    const double *Y;
    const double *Xn = nullptr;  // subtracted in the residual norm (FP64: cur)
    if (!mixed) {
      // nxt = E + F cur; the residual of cur is nxt - cur
      nparts = spmm<double, double>(s, ctx, F, S, F->d_rval, cur(k), Er, nxt(k), 0, done,
                                    SweepOut{nullptr, nullptr, part, 1}, "radiosity_fp64_sweep");
      Y = nxt(k);
      Xn = cur(k);
    } else {
      // d = D_k: B += d, r <- (r - d) + F d (FP64 values and accumulation),
      // D_{k+1} = fp32(r)
      float *Dk = k & 1 ? D1 : D0, *Dk1 = k & 1 ? D0 : D1;
      nparts = spmm<double, float>(s, ctx, F, S, F->d_rval, Dk, rk(k), rk1(k), 1, done,
                                   SweepOut{Bacc, Dk1, part, 0}, "radiosity_mixed_sweep");
      Y = rk1(k);
    }
    if (!nparts) {
      colnorm_partial_kernel<double><<<G, 256, 0, s>>>(nf, S, Y, Xn, part, done);
      ILLUM_LAUNCH_CHECK();
      nparts = G;
    }
    colnorm_final_kernel<<<fb, 256, 0, s>>>(nparts, S, part, col, e2, ctl, hist, rel_tol, 1);
    ILLUM_LAUNCH_CHECK();
    if (fb > 1) {
      decide_kernel<<<1, 256, 0, s>>>(S, col, e2, ctl, hist, rel_tol);
      ILLUM_LAUNCH_CHECK();
    }
  };

  // Sweeps are launched in batches (the first as long as the previous solve
  // of this matrix took); the control block and the residual history are
  // read back, in one copy, after each.
  std::vector<char> hbuf(o_hist + 8 * ((size_t)max_iter + 1));
  SolveCtl h{};
  int launched = 0;
  int batch = F->last_nprod > 0 ? F->last_nprod : 4;
  for (;;) {
    const int nb = std::min(batch, max_iter - launched);
    for (int j = 0; j < nb; ++j) sweep(launched++);
    ILLUM_HIP(hipMemcpyAsync(hbuf.data(), ws, o_hist + 8 * (size_t)launched,
                             hipMemcpyDeviceToHost, s));
    sync(ctx);
    std::memcpy(&h, hbuf.data(), sizeof(SolveCtl));
    if (h.done || launched >= max_iter) break;
    batch = std::min(2 * batch, 64);
  }
  const int nprod = h.nprod;
  F->last_nprod = nprod;
  std::vector<double> hh(nprod);
  std::memcpy(hh.data(), hbuf.data() + o_hist, 8 * (size_t)nprod);
  // FP64: converged at sweep k -> B_k = cur(k); otherwise the newest iterate.
  const double *Bcur = mixed ? Bacc : cur(h.done ? nprod - 1 : nprod);

  // Output to column-major.
  DBuf<double> Bout;
  double *dst = B;
  if (!out_dev) { Bout.reset(n); dst = Bout.p; }
  if (S == 1) {
    if (Bcur != dst) ILLUM_HIP(hipMemcpyAsync(dst, Bcur, 8 * n, hipMemcpyDeviceToDevice, s));
  } else {
    transpose(s, nf, S, Bcur, dst, false);
  }
  if (!out_dev) ILLUM_HIP(hipMemcpyAsync(B, dst, 8 * n, hipMemcpyDeviceToHost, s));
  if (!(in_dev && out_dev && (flags & ILLUM_ASYNC))) sync(ctx);
  if (resid) std::copy(hh.begin(), hh.end(), resid);
  if (iters) *iters = nprod;
  const double rel = nprod > 0 ? hh[nprod - 1] : 1.0;
  if (!(rel <= rel_tol))
    throw Error(ILLUM_ERR_NOT_CONVERGED,
                "radiosity solve did not reach the tolerance (relative residual " +
                    std::to_string(rel) + ")");
}

}  // namespace illum
