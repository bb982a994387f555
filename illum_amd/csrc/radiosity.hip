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
//   ILLUM_RAD_MIXED: iterative refinement with one-sweep corrections.  The
//     FP64 residual r = E - (I-F)B is carried from sweep to sweep.  Each
//     sweep rounds it to a correction d = fp16(r 2^e), with a per-column
//     power of two 2^e that puts the column's 2-norm at or below 2^13, and
//     adds delta = d 2^-e (exact in FP64) to B.  It updates
//     r <- (r - delta) + F delta with FP64 values and accumulation, gathering
//     d in FP16.  So r stays the residual of the B built so far: rounding d
//     only slows the fall of r (by about 2^-11 per sweep), never its
//     accuracy.
// Both report max_s ||E_s - (I-F)B_s||_2 / ||E_s||_2, measured in FP64.
#include "hip_util.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

namespace illum {

// Scratch of illum_radiosity_solve_noalloc: one fixed array per GPU (each
// device that loads the library holds its own copy until the process ends).
// Host code never names the array in a kernel argument or a copy: it uses the
// device address that radiosity_noalloc_scratch resolved for that device.
constexpr int RAD_NOALLOC_MAX_ITER = 256;
static __device__ __align__(256) unsigned char scratch[RAD_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[MAX_DEVICES];

// Device address of `scratch` on `device`, which must be the current device
// the first time it is asked for (illum_ctx_create calls this at set-up).
void *radiosity_noalloc_scratch(int device) {
  if (device < 0 || device >= MAX_DEVICES) throw Error(ILLUM_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    ILLUM_HIP(hipGetDevice(&cur));
    if (cur != device) throw Error(ILLUM_ERR_ARG, "no-allocation scratch resolved off its device");
    void *p = nullptr;
    ILLUM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(ILLUM_ERR_HIP, "hipGetSymbolAddress returned no address");
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

// Where a solve's device scratch comes from: the context's grow-only
// workspace slots, or (fixed != nullptr) consecutive 256-byte aligned pieces
// of a fixed array of `cap` bytes, which allocates nothing.
struct RadScratch {
  illum_ctx *ctx;
  unsigned char *fixed;
  size_t cap, used;
  void *get(int slot, size_t bytes) {
    if (!fixed) return ctx_workspace(ctx, slot, bytes);
    const size_t off = (used + 255) & ~(size_t)255;
    if (off > cap || bytes > cap - off)
      throw Error(ILLUM_ERR_CAPACITY,
                  "radiosity solve needs more than the " + std::to_string(cap >> 20) +
                      " MiB no-allocation scratch (nf x S too large)");
    used = off + bytes;
    return fixed + off;
  }
};

// Scale exponent of the mixed method's FP16 corrections for a column whose
// residual has squared 2-norm nrm2: d = fp16(r 2^e), ||r||_2 2^e <= 2^13.
// A column's next residual is bounded element-wise by ||F||_inf times this
// norm, so elements stay below FP16's 65504 while ||F||_inf < 8 (row sums of
// F are at most the albedo for physical form factors).
__device__ __forceinline__ int half_scale_exp(double nrm2) {
  const double nrm = sqrt(nrm2);
  if (!(nrm > 0.0) || nrm == (double)INFINITY) return 0;  // zero, NaN: any scale
  int q;
  (void)frexp(nrm, &q);  // nrm < 2^q
  return max(-1000, min(1000, 13 - q));
}

// fp16(y scale), clamped to FP16's finite range: a clamped element leaves a
// larger remainder in the residual, which the next sweeps remove.
__device__ __forceinline__ _Float16 to_half_scaled(double y, double scale) {
  const float v = (float)(y * scale);
  return (_Float16)fminf(fmaxf(v, -65504.0f), 65504.0f);
}

// Sparse products over the CSR rows of F (FP64 values) with a gathered block
// X and right-hand side R / output Y (FP64 accumulation).  Rows of X, R and Y
// hold S contiguous values (row-major [nf][S]).
//   FP64 Jacobi (X FP64):  Y = R + F X; the residual element is Y - X.
//   mixed (X = d_k, FP16): with delta_k = d_k 2^-ek[s] at the row,
//                          Y = r_{k+1} = (R - delta_k) + (F d_k) 2^-ek[s]
//                          (R = r_k), Dn = d_{k+1} = fp16(Y 2^en[s]), and in
//                          B mode B += delta_k; the residual element is Y.
// The rows and wide kernels sum a row's nonzeros in CSR order (identical
// results whatever the lane width); the small-S kernel uses a fixed
// lane-strided order and tree.  Rows go to XCDs in contiguous ranges:
// neighbouring facets share most of their visible set, so gathered rows of X
// are reused from the XCD's L2.  Each kernel does the per-element work of the
// sweep on the rows it owns (no other launch touches the vectors in between).
struct MixedIO {
  const int *ek;  // exponents of the gathered correction d_k
  const int *en;  // exponents of the next correction d_{k+1}
  _Float16 *Dn;   // d_{k+1}
  double *B;      // B mode: B += delta_k; null: the corrections stay in a ring
  int xc;         // row groups per chunk dealt to the XCDs in turn (0: one range per XCD)
  // Ring mode's FP32 remainders (the wide kernel): eps_k = r_k - delta_k is
  // kept as q_k = fp32(eps_k 2^(ek + 11)) instead of r_k in FP64.  Qk null:
  // eps_k is formed from R = r_k and d_k (r_k in FP64: the first product,
  // r_0 = E, or every product without remainders).  Qn non-null receives
  // q_{k+1} (Y is not written) and the column sums of eps_{k+1}^2 go to part
  // after the residuals'; null: Y receives r_{k+1} in FP64.
  const float *Qk;
  float *Qn;
};

__device__ __forceinline__ int64_t sweep_block(int b, int n, int xc) {
  return xc > 0 ? xcd_chunked(b, n, xc) : xcd_remap(b, n);
}

template <class TX>
constexpr bool is_mixed() { return std::is_same<TX, _Float16>::value; }

// Wide form (S >= SWEEP_WIDE_MIN_S): lane l covers PER consecutive
// right-hand sides (one vector load of PER values per nonzero); with S < 64
// the lanes past S run the gathers on column 0 and discard them.  The grid is column-chunk-major:
// chunk ch (columns [64 PER ch, 64 PER (ch + 1))) runs over all row blocks
// before the next chunk starts.  Per chunk there are nblk row blocks (a
// multiple of 8); block g owns rows [rblk[g], rblk[g + 1]), contiguous and
// of about equal weight (nonzeros plus SWEEP_ROW_WEIGHT per row, see
// row_blocks_kernel), interleaved over its four waves (wave w takes rows
// rblk[g] + w + 4 i).  With one contiguous range of blocks per XCD, each
// XCD's rows are a compact patch of the surface (facet order is spatially
// coherent), so the rows of X they gather stay in that XCD's L2; weighting
// by nonzeros keeps the XCDs' shares even where crater floors have rows of
// hundreds of entries.  Requires S % PER == 0.
// The block's column sums of squares of the residual elements go to
// part[g][S]: each wave sums its rows in order, then waves 0..3 in order,
// so the sums do not depend on PER or on where a block runs.
constexpr int SWEEP_WIDE_MIN_S = 32;  // right-hand sides from which products take the wide form
constexpr int RAD_RING_MIN_S = 128;   // mixed: corrections in a ring from here, else B mode
constexpr int SWEEP_RPW = 4;         // rows per wave of a row block, on average
constexpr int SWEEP_ROW_WEIGHT = 8;  // a row's fixed cost, in nonzeros
constexpr int SWEEP_UNROLL = 8;      // gathers in flight per wave

// Row blocks of the wide product: rblk[k] = the largest m in [0, nf] with
// rowptr[m] + SWEEP_ROW_WEIGHT m <= k W, W = total weight / nblk, for
// k = 0..nblk (rblk[0] = 0, rblk[nblk] = nf).  Monotone in k.
__global__ __launch_bounds__(256) void row_blocks_kernel(int64_t nf,
                                                          const int64_t *__restrict__ rowptr,
                                                          int nblk, int *__restrict__ rblk) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k > nblk) return;
  if (k == nblk) { rblk[k] = (int)nf; return; }
  const int64_t tot = rowptr[nf] + SWEEP_ROW_WEIGHT * nf;
  // floor(tot k / nblk) without forming tot k: (tot mod nblk) k < nblk^2 <= 2^56
  const int64_t T = (tot / nblk) * k + (tot % nblk) * k / nblk;
  int64_t lo = 0, hi = nf;  // c(lo) <= T, find the largest such m
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (rowptr[mid] + SWEEP_ROW_WEIGHT * mid <= T) lo = mid; else hi = mid - 1;
  }
  rblk[k] = (int)lo;
}

// acc[p] += f x[p] with one rounding (a fused multiply-add per element).
template <int PER, class AV, class XV>
__device__ __forceinline__ void fma_acc(AV &acc, double f, const XV &x) {
#pragma unroll
  for (int p = 0; p < PER; ++p) acc[p] = __builtin_fma(f, (double)x[p], acc[p]);
}

// Gathers of rows of X: `off` is the byte offset of the row (a wave-uniform
// value read from a lane), `lo` the lane's byte offset within it.  BUF:
// buffer loads with the row offset as the scalar offset (no per-lane address
// arithmetic; the block must span < 2 GiB); else global loads.
template <class XV, bool BUF>
struct Gather {
  const char *base;
  __amdgpu_buffer_rsrc_t rs;
  unsigned lo;
  __device__ Gather(const void *X, unsigned xbytes, unsigned lane_off) : base((const char *)X), lo(lane_off) {
    if constexpr (BUF) rs = __builtin_amdgcn_make_buffer_rsrc((void *)X, 0, (int)xbytes, 0x00020000);
  }
  __device__ __forceinline__ XV operator()(unsigned off) const {
    if constexpr (BUF) {
      if constexpr (sizeof(XV) == 4)
        return __builtin_bit_cast(XV, __builtin_amdgcn_raw_buffer_load_b32(rs, lo, off, 0));
      else if constexpr (sizeof(XV) == 8)
        return __builtin_bit_cast(XV, __builtin_amdgcn_raw_buffer_load_b64(rs, lo, off, 0));
      else
        return __builtin_bit_cast(XV, __builtin_amdgcn_raw_buffer_load_b128(rs, lo, off, 0));
    } else {
      return *(const XV *)(base + (size_t)off + lo);
    }
  }
};

// A row's nonzeros are taken 64 at a time: the whole wave loads 64 column
// indices (one per lane), turns them into byte offsets of rows of X, and the
// gathers read the offsets back with readlane, so the SWEEP_UNROLL gathers
// of a group issue back to back; the values come by scalar loads.  Each
// element is accumulated with fused multiply-adds in CSR order.
template <class TX, int PER, bool BUF>
__global__ __launch_bounds__(256) void sweep_wide_kernel(
    int64_t nf, int S, int nblk, const int *__restrict__ rblk, const int64_t *__restrict__ rowptr,
    const int32_t *__restrict__ colidx, const double *__restrict__ val, const TX *__restrict__ X,
    unsigned xbytes, const double *__restrict__ R, double *__restrict__ Y,
    const int *__restrict__ done, double *__restrict__ part, MixedIO m) {
  if (*done) return;  // uniform: the whole grid returns or none of it
  constexpr int U = SWEEP_UNROLL;
  typedef TX xvec __attribute__((ext_vector_type(PER)));
  typedef double yvec __attribute__((ext_vector_type(PER)));
  typedef float fvec __attribute__((ext_vector_type(PER)));
  static_assert(sizeof(xvec) == 4 || sizeof(xvec) == 8 || sizeof(xvec) == 16 || !BUF,
                "buffer gathers of 4, 8 or 16 bytes");
  __shared__ yvec red[3][64], rede[3][64];
  const int ch = blockIdx.x / nblk, gb = blockIdx.x - ch * nblk;
  const int64_t g = sweep_block(gb, nblk, m.xc);
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const int s = (ch * 64 + lane) * PER;
  const bool on = s < S;
  // Every lane runs the gathers (readlane needs the index lanes active); a
  // lane past S reads column 0 of the same rows and discards the sums.
  const Gather<xvec, BUF> gather(X, xbytes, (unsigned)(on ? s : 0) * (unsigned)sizeof(TX));
  const unsigned rowbytes = (unsigned)S * (unsigned)sizeof(TX);
  yvec sq = (yvec)0.0, se = (yvec)0.0;  // column sums of r_{k+1}^2 and eps_{k+1}^2
  const int64_t r1 = rblk[g + 1];
  for (int64_t row = rblk[g] + w; row < r1; row += 4) {
    const int64_t k0 = rowptr[row], k1 = rowptr[row + 1];
    const int64_t at = row * S + s;
    yvec acc = (yvec)0.0;
    for (int64_t kb = k0; kb < k1; kb += 64) {
      const int cnt = (int)std::min<int64_t>(64, k1 - kb);
      // past the row: another valid index or the zero pad
      const unsigned bo = (unsigned)colidx[kb + lane] * rowbytes;
      const double *vk = val + kb;  // wave-uniform: scalar loads
      int j = 0;
      for (; j + U <= cnt; j += U) {
        xvec x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = gather((unsigned)__builtin_amdgcn_readlane((int)bo, j + u));
#pragma unroll
        for (int u = 0; u < U; ++u) fma_acc<PER>(acc, vk[j + u], x[u]);
      }
      for (; j < cnt; ++j)
        fma_acc<PER>(acc, vk[j], gather((unsigned)__builtin_amdgcn_readlane((int)bo, j)));
    }
    if (!on) continue;
    if constexpr (is_mixed<TX>()) {
      yvec ik, sn, isn;
#pragma unroll
      for (int p = 0; p < PER; ++p) {
        ik[p] = ldexp(1.0, -m.ek[s + p]);
        sn[p] = ldexp(1.0, m.en[s + p]);
        isn[p] = ldexp(1.0, -m.en[s + p]);
      }
      // eps_k = r_k - delta_k, exact in FP64 (delta_k is r_k rounded)
      yvec eps;
      if (m.Qk) {
        const fvec q = *(const fvec *)(m.Qk + at);
#pragma unroll
        for (int p = 0; p < PER; ++p) eps[p] = (double)q[p] * (ik[p] * 0x1p-11);
      } else {
        const yvec xs = __builtin_convertvector(*(const xvec *)(X + at), yvec) * ik;  // delta_k
        eps = *(const yvec *)(R + at) - xs;
        if (m.B) *(yvec *)(m.B + at) = *(const yvec *)(m.B + at) + xs;  // B mode
      }
      const yvec y = eps + acc * ik;  // r_{k+1}
      sq = sq + y * y;
      xvec dn;
#pragma unroll
      for (int p = 0; p < PER; ++p) dn[p] = to_half_scaled(y[p], sn[p]);
      *(xvec *)(m.Dn + at) = dn;
      if (m.Qn) {
        // eps_{k+1} = r_{k+1} - delta_{k+1} (exact), stored rounded to FP32
        yvec en;
        fvec qn;
#pragma unroll
        for (int p = 0; p < PER; ++p) {
          en[p] = y[p] - (double)dn[p] * isn[p];
          qn[p] = (float)(en[p] * (sn[p] * 0x1p11));
        }
        se = se + en * en;
        *(fvec *)(m.Qn + at) = qn;
      } else {
        *(yvec *)(Y + at) = y;
      }
    } else {
      const yvec x = __builtin_convertvector(*(const xvec *)(X + at), yvec);
      const yvec y = *(const yvec *)(R + at) + acc;
      *(yvec *)(Y + at) = y;
      const yvec d = y - x;
      sq = sq + d * d;
    }
  }
  if (w > 0) red[w - 1][lane] = sq;
  if (m.Qn && w > 0) rede[w - 1][lane] = se;
  __syncthreads();
  if (w == 0 && on) {
    yvec t = sq + red[0][lane];
    t = t + red[1][lane];
    t = t + red[2][lane];
    *(yvec *)(part + g * S + s) = t;
    if (m.Qn) {
      // the remainders' column sums follow the residuals' (nblk rows later)
      yvec te = se + rede[0][lane];
      te = te + rede[1][lane];
      te = te + rede[2][lane];
      *(yvec *)(part + ((int64_t)nblk + g) * S + s) = te;
    }
  }
}

// Small S (< 32): LPR lanes per row (64 / LPR rows per wave), lanes strided
// over the row's nonzeros, fixed-shape reduction per right-hand side; the
// block's column sums of squares of the residual elements go to
// part[block][S] (rows in order).
constexpr int SPMM_NNZ_LPR = 16;
constexpr int SPMM_NNZ_ROWS = 4 * 64 / SPMM_NNZ_LPR;  // rows per 256-thread block
constexpr int SPMM_NNZ_MAXS = 31;

template <class TX>
__global__ __launch_bounds__(256) void sweep_nnz_kernel(
    int64_t nf, int S, const int64_t *__restrict__ rowptr, const int32_t *__restrict__ colidx,
    const double *__restrict__ val, const TX *__restrict__ X, const double *__restrict__ R,
    double *__restrict__ Y, const int *__restrict__ done, double *__restrict__ part, MixedIO m) {
  constexpr int LPR = SPMM_NNZ_LPR;
  __shared__ double sq[SPMM_NNZ_ROWS][SPMM_NNZ_MAXS + 1];
  if (*done) return;  // uniform: the whole grid returns or none of it
  const int rib = threadIdx.x / LPR;  // row in block
  const int sub = threadIdx.x % LPR;
  const int64_t row = sweep_block(blockIdx.x, gridDim.x, m.xc) * SPMM_NNZ_ROWS + rib;
  const bool valid = row < nf;
  const int64_t k0 = valid ? rowptr[row] : 0, k1 = valid ? rowptr[row + 1] : 0;
  for (int s = 0; s < S; ++s) {
    // the row's own operands are loaded ahead of the gather
    const int64_t at = row * S + s;
    const bool own = valid && sub == 0;
    const double r_at = own ? R[at] : 0.0, x_at = own ? (double)X[at] : 0.0;
    const double b_at = own && m.B ? m.B[at] : 0.0;
    double acc = 0;
#pragma unroll 4
    for (int64_t k = k0 + sub; k < k1; k += LPR)
      acc = __builtin_fma(val[k], (double)X[(int64_t)colidx[k] * S + s], acc);
    for (int w = LPR / 2; w > 0; w >>= 1) acc = acc + __shfl_xor(acc, w);
    if (sub == 0) {
      double d2 = 0.0;
      if (valid) {
        if constexpr (is_mixed<TX>()) {
          const double ik = ldexp(1.0, -m.ek[s]);
          const double x = x_at * ik;
          const double y = (r_at - x) + acc * ik;
          Y[at] = y;
          if (m.B) m.B[at] = b_at + x;
          m.Dn[at] = to_half_scaled(y, ldexp(1.0, m.en[s]));
          d2 = y * y;
        } else {
          const double y = r_at + acc;
          Y[at] = y;
          const double d = y - x_at;
          d2 = d * d;
        }
      }
      sq[rib][s] = d2;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < S) {
    double t = 0.0;
    for (int r = 0; r < SPMM_NNZ_ROWS; ++r) t += sq[r][threadIdx.x];
    part[(int64_t)blockIdx.x * S + threadIdx.x] = t;
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
  int floor;  // FP32 remainders: their rounding bound reached the tolerance (done is set too)
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
#pragma unroll 16
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

// Bound on the FP32 rounding of a stored remainder q = fp32(eps 2^(e + 11)),
// unscaled, in the 2-norm over nf elements: 2^-24 ||eps|| for normal FP32
// values plus 2^-150 2^-(e + 11) per element for subnormal ones (taken twice
// as large).  en = ||eps||_2, summed in FP64 before the rounding (itself
// rounded by at most a few ulps, which the factor two covers).
__device__ __forceinline__ double remainder_bound(double en, int e, double sqrt_nf) {
  return 0x1p-24 * en * 1.0000001 + sqrt_nf * ldexp(1.0, -(160 + e));
}

// The stop decision: rel = max_s sqrt(col_s / e2_s) (sqrt(col_s) where e2_s =
// 0) is stored as hist[nprod], the product counted, done set when rel <= tol.
// ex (nullable, the mixed method): col is the residual r_{k+1} of product k =
// nprod, and ex[k + 2][s] receives the scale exponent of d_{k+2}, which
// product k + 1 writes (r_{k+2} is bounded by r_{k+1}).
// tb (nullable, FP32 remainders): the carried r_{k+1} is the residual of
// B_{k+1} up to the FP32 roundings t_i of the stored remainders q_1..q_k, so
// ||E - (I - F) B_{k+1}|| <= ||r_{k+1}|| + tb_s with tb_s the sum of their
// bounds; rel uses that sum, and tb_s then takes the bound of q_{k+1} (from
// cole_s = ||eps_{k+1}||^2, scaled by ex[k + 1][s]), which product k wrote.
// Once tb_s >= tol ||E_s|| no later product can meet the tolerance: floor
// and done are set (the host solves again with FP64 residuals).
// Called by one whole block after col[] is complete.
__device__ void decide(int S, const double *col, const double *e2, SolveCtl *ctl, double *hist,
                       double tol, double *red, int *ex, double *tb, const double *cole,
                       double sqrt_nf) {
  const int k = ctl->nprod;
  double m = 0.0;
  int fl = 0;
  for (int s = threadIdx.x; s < S; s += 256) {
    const double r2 = col[s];
    if (tb) {
      const double en = sqrt(e2[s]);
      const double cert = sqrt(r2) + tb[s];
      m = max_nan(m, en > 0 ? cert / en : cert);
      const double t = tb[s] + remainder_bound(sqrt(cole[s]), ex[(int64_t)(k + 1) * S + s], sqrt_nf);
      tb[s] = t;
      if (!(t < tol * (en > 0 ? en : 1.0))) fl = 1;
    } else {
      m = max_nan(m, e2[s] > 0 ? sqrt(r2 / e2[s]) : sqrt(r2));
    }
    if (ex) ex[(int64_t)(k + 2) * S + s] = half_scale_exp(r2);
  }
  fl = __syncthreads_or(fl);  // block-wide: some column's bound reached the tolerance
  __syncthreads();
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = max_nan(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double rel = red[0];
    hist[k] = rel;
    ctl->nprod = k + 1;
    if (fl) ctl->floor = 1;
    if (rel <= tol || fl) ctl->done = 1;
  }
}

// Adds the G partials of columns [b cw, (b + 1) cw) (block b) in a fixed
// order into col[].  record = 1: nothing once done is set, and a single block
// also takes the stop decision (with several blocks decide_kernel follows).
// First stage of the wide product's column sums: its G partial rows (two
// sets of G with FP32 remainders: blockIdx.z) reduced to WIDE_SUM_P rows per
// set, out[set][p][s].  Block (c, p, set) sums columns [64 c, 64 c + 64) over
// row chunk p: wave w takes rows w, w + 4, ... of the chunk in order (a
// coalesced 512-byte read per row), then the waves are added in order 0..3.
constexpr int WIDE_SUM_P = 64;

__global__ __launch_bounds__(256) void wide_colsum_kernel(int G, int S,
                                                          const double *__restrict__ part,
                                                          double *__restrict__ out,
                                                          const int *__restrict__ done) {
  if (*done) return;
  __shared__ double red[4][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = blockIdx.x * 64 + lane, p = blockIdx.y, set = blockIdx.z;
  const int per = (G + WIDE_SUM_P - 1) / WIDE_SUM_P;
  const int r0 = std::min(G, p * per), r1 = std::min(G, r0 + per);
  const double *in = part + (int64_t)set * G * S;
  double acc = 0.0;
  if (s < S)
    for (int r = r0 + w; r < r1; r += 4) acc += in[(int64_t)r * S + s];
  red[w][lane] = acc;
  __syncthreads();
  if (w == 0 && s < S) {
    double t = red[0][lane] + red[1][lane];
    t = t + red[2][lane];
    t = t + red[3][lane];
    out[((int64_t)set * WIDE_SUM_P + p) * S + s] = t;
  }
}

// record = 0 (the solve's start, col = ||E_s||^2) also resets the control
// block (and tb, when given), and with ex (the mixed method) ex[0][s] and
// ex[1][s] receive the scale exponents of d_0 and d_1.
constexpr int COLNORM_CW = 32;

__global__ __launch_bounds__(256) void colnorm_final_kernel(int G, int S, int cw,
                                                            const double *__restrict__ part,
                                                            double *__restrict__ col,
                                                            const double *__restrict__ e2,
                                                            SolveCtl *__restrict__ ctl,
                                                            double *__restrict__ hist, double tol,
                                                            int record, int *__restrict__ ex,
                                                            double *__restrict__ tb,
                                                            double *__restrict__ cole,
                                                            double sqrt_nf) {
  if (record && ctl->done) return;
  if (!record && blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->done = 0;
    ctl->nprod = 0;
    ctl->floor = 0;
  }
  __shared__ double red[256];
  const int c0 = blockIdx.x * cw;
  const int cwb = std::min(cw, S - c0);
  block_colsum<false>(0, G, S, c0, cwb, part, (const double *)nullptr, red, col);
  // cole (record, FP32 remainders): the remainders' sums, G rows after the residuals'
  if (record && cole)
    block_colsum<false>(0, G, S, c0, cwb, part + (int64_t)G * S, (const double *)nullptr, red,
                        cole);
  if (record && gridDim.x == 1) {
    __syncthreads();
    decide(S, col, e2, ctl, hist, tol, red, ex, tb, cole, sqrt_nf);
  } else if (!record && (int)threadIdx.x < cwb) {
    // thread t wrote col[c0 + t] itself
    if (ex) {
      const int e = half_scale_exp(col[c0 + threadIdx.x]);
      ex[c0 + threadIdx.x] = e;
      ex[S + c0 + threadIdx.x] = e;
    }
    if (tb) tb[c0 + threadIdx.x] = 0.0;
  }
}

// One right-hand side without FP32 remainders (S = 1, B mode): the G
// partials summed by 1,024 threads -- each a strided run in order, then a
// fixed shuffle tree per wavefront and the 16 wavefront sums in order -- and
// the stop decision taken by thread 0, as decide does for one column.  Four
// times the threads of colnorm_final_kernel and no block-wide reductions in
// the decision: the small solves of a time loop take one after every product.
__global__ __launch_bounds__(1024) void colnorm_final1_kernel(int G,
                                                              const double *__restrict__ part,
                                                              double *__restrict__ col,
                                                              const double *__restrict__ e2,
                                                              SolveCtl *__restrict__ ctl,
                                                              double *__restrict__ hist,
                                                              double tol, int *__restrict__ ex) {
  if (ctl->done) return;
  __shared__ double wsum[16];
  double acc = 0.0;
  for (int i = threadIdx.x; i < G; i += 1024) acc += part[i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double r2 = 0.0;
  for (int w = 0; w < 16; ++w) r2 += wsum[w];
  col[0] = r2;
  const int k = ctl->nprod;
  const double rel = max_nan(0.0, e2[0] > 0 ? sqrt(r2 / e2[0]) : sqrt(r2));
  if (ex) ex[k + 2] = half_scale_exp(r2);
  hist[k] = rel;
  ctl->nprod = k + 1;
  if (rel <= tol) ctl->done = 1;
}

__global__ __launch_bounds__(256) void decide_kernel(int S, const double *__restrict__ col,
                                                     const double *__restrict__ e2,
                                                     SolveCtl *__restrict__ ctl,
                                                     double *__restrict__ hist, double tol,
                                                     int *__restrict__ ex, double *__restrict__ tb,
                                                     const double *__restrict__ cole,
                                                     double sqrt_nf) {
  if (ctl->done) return;
  __shared__ double red[256];
  decide(S, col, e2, ctl, hist, tol, red, ex, tb, cole, sqrt_nf);
}

// Queue-only solves (a time loop's steps): after the fixed product budget,
// one thread folds the solve's outcome into the loop's running record.
__global__ void solve_stats_kernel(const SolveCtl *__restrict__ ctl,
                                   const double *__restrict__ hist,
                                   SolveStats *__restrict__ st) {
  const int k = ctl->nprod;
  const double rel = k > 0 ? hist[k - 1] : 1.0;
  st->solves += 1;
  st->max_products = max(st->max_products, k);
  if (!ctl->done) st->unconverged += 1;
  st->max_resid = max_nan(st->max_resid, rel);
}

// Mixed method start (row-major E): d_0 = fp16(E 2^ex[0][s]), B = 0 (B
// nullable).  r_0 = E is read from E itself by the first product.
__global__ void mixed_init_kernel(int64_t n, int S, const double *__restrict__ E,
                                  const int *__restrict__ ex, _Float16 *__restrict__ D,
                                  double *__restrict__ B) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  D[k] = to_half_scaled(E[k], ldexp(1.0, ex[k % S]));
  if (B) B[k] = 0.0;
}

// Ring mode (S >= 128): the corrections d_j stay in ring slot j % R and B is
// their sum.  fold_kernel adds delta_j for j in [a, min(b, nprod)) to Bacc in
// order of j (from zero when init), before product b overwrites slot
// (b + 1) % R; it does nothing for sweeps that never ran.
__global__ __launch_bounds__(256) void fold_kernel(int64_t n, int S, int R,
                                                   const _Float16 *__restrict__ ring,
                                                   const int *__restrict__ ex,
                                                   const SolveCtl *__restrict__ ctl, int a, int b,
                                                   int init, double *__restrict__ Bacc) {
  const int j1 = min(b, ctl->nprod);
  if (a >= j1) return;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * 256) {
    const int s = (int)(k % S);
    double acc = init ? 0.0 : Bacc[k];
    for (int j = a; j < j1; ++j)
      acc += (double)ring[(int64_t)(j % R) * n + k] * ldexp(1.0, -ex[(int64_t)j * S + s]);
    Bacc[k] = acc;
  }
}

// Ring mode output: out (column-major nf x S) = [Bacc +] sum of delta_j over
// j in [min(folded, nprod), nprod), in order of j (nprod read on the device,
// so the output is queued before the host reads the control block).  A
// block takes 32 rows x 64 columns: with
// S % 4 == 0 a thread reads 4 consecutive values of a row per slot (16
// threads per row, two passes of 16 rows), else one value at a time; the
// tile is written transposed, 32 consecutive rows per column.
__global__ __launch_bounds__(256) void ring_out_kernel(int64_t nf, int S, int R,
                                                       const _Float16 *__restrict__ ring,
                                                       const int *__restrict__ ex, int folded,
                                                       const SolveCtl *__restrict__ ctl,
                                                       const double *__restrict__ Bacc,
                                                       double *__restrict__ out) {
  __shared__ double tile[32][65];
  const int j1 = ctl->nprod, j0 = min(folded, j1);
  const int64_t n = nf * S;
  const int64_t i0 = (int64_t)blockIdx.x * 32;
  const int s0 = blockIdx.y * 64;
  const int t = threadIdx.x;
  if (S % 4 == 0) {
    typedef _Float16 h4 __attribute__((ext_vector_type(4)));
    const int c = (t & 15) * 4;  // columns s0 + c .. + 3
    const int s = s0 + c;
    for (int k = t >> 4; k < 32; k += 16) {
      const int64_t i = i0 + k;
      if (i >= nf || s >= S) continue;
      const int64_t at = i * S + s;
      double v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = Bacc ? Bacc[at + q] : 0.0;
      for (int j = j0; j < j1; ++j) {
        const h4 d = *(const h4 *)(ring + (int64_t)(j % R) * n + at);
        const int *e = ex + (int64_t)j * S + s;
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] += (double)d[q] * ldexp(1.0, -e[q]);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) tile[k][c + q] = v[q];
    }
  } else {
    const int c = t & 63;
    const int s = s0 + c;
    for (int k = t >> 6; k < 32; k += 4) {
      const int64_t i = i0 + k;
      if (i >= nf || s >= S) continue;
      const int64_t at = i * S + s;
      double v = Bacc ? Bacc[at] : 0.0;
      for (int j = j0; j < j1; ++j)
        v += (double)ring[(int64_t)(j % R) * n + at] * ldexp(1.0, -ex[(int64_t)j * S + s]);
      tile[k][c] = v;
    }
  }
  __syncthreads();
  const int i = t & 31;
  for (int c = t >> 5; c < 64; c += 8) {
    const int s = s0 + c;
    if (i0 + i < nf && s < S) out[(int64_t)s * nf + i0 + i] = tile[i][c];
  }
}

// Column sums of squares of a column-major (nf x S) array: block (s, p)
// sums rows [p rows_per, (p + 1) rows_per) of column s in a fixed order into
// part[p][s] (colnorm_final_kernel adds the partials).
__global__ __launch_bounds__(256) void colnorm_cm_partial_kernel(int64_t nf, int S,
                                                                 int64_t rows_per,
                                                                 const double *__restrict__ A,
                                                                 double *__restrict__ part) {
  __shared__ double red[256];
  const int s = blockIdx.x, p = blockIdx.y;
  const int64_t r0 = std::min<int64_t>(nf, p * rows_per), r1 = std::min<int64_t>(nf, r0 + rows_per);
  const double *c = A + (int64_t)s * nf;
  double acc = 0.0;
#pragma unroll 4
  for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) {
    const double v = c[i];
    acc += v * v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)p * S + s] = red[0];
}

// Right-hand sides to row-major: Er = E (column-major nf x S in), and for
// the mixed method d_0 = fp16(E 2^ex[0][s]) (D0 nullable) and B = 0 (Bacc
// nullable), through a 32 x 32 tile transpose.
__global__ __launch_bounds__(256) void load_rhs_kernel(int64_t nf, int S,
                                                       const double *__restrict__ E,
                                                       double *__restrict__ Er,
                                                       const int *__restrict__ ex,
                                                       _Float16 *__restrict__ D0,
                                                       double *__restrict__ Bacc) {
  __shared__ double tile[32][33];
  const int64_t i0 = (int64_t)blockIdx.x * 32;
  const int s0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int64_t i = i0 + tx;
    const int s = s0 + k;
    if (i < nf && s < S) tile[k][tx] = E[(int64_t)s * nf + i];
  }
  __syncthreads();
  const int s = s0 + tx;
  const double scale = D0 && s < S ? ldexp(1.0, ex[s]) : 1.0;
  for (int k = ty; k < 32; k += 8) {
    const int64_t i = i0 + k;
    if (i < nf && s < S) {
      const double v = tile[tx][k];
      const int64_t at = i * S + s;
      Er[at] = v;
      if (D0) D0[at] = to_half_scaled(v, scale);
      if (Bacc) Bacc[at] = 0.0;
    }
  }
}

// FP64 output: B = cur(k), k = nprod - 1 once done, else nprod (the newest
// iterate), where cur(0) = E and then X1, X0, X1, ... (row-major in,
// column-major out).  The source is chosen on the device, so the output is
// queued before the host reads the control block.
__global__ __launch_bounds__(256) void fp64_out_kernel(int64_t nf, int S,
                                                       const double *__restrict__ Er,
                                                       const double *__restrict__ X0,
                                                       const double *__restrict__ X1,
                                                       const SolveCtl *__restrict__ ctl,
                                                       double *out) {
  __shared__ double tile[32][33];
  const int k = ctl->done ? ctl->nprod - 1 : ctl->nprod;
  const double *in = k == 0 ? Er : (k & 1 ? X1 : X0);
  const int64_t i0 = (int64_t)blockIdx.x * 32;
  const int s0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int q = ty; q < 32; q += 8) {
    const int64_t i = i0 + q;
    const int s = s0 + tx;
    if (i < nf && s < S) tile[q][tx] = in[i * S + s];
  }
  __syncthreads();  // (with S = 1 and out == Er a block rewrites only what it read)
  for (int q = ty; q < 32; q += 8) {
    const int64_t i = i0 + tx;
    const int s = s0 + q;
    if (i < nf && s < S) out[(int64_t)s * nf + i] = tile[tx][q];
  }
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

// Sparse product of one sweep, dispatched by S (see the kernels above).
// Returns the number of column-sum partial rows written to part (0: none,
// the rows kernel; the caller then runs colnorm_partial_kernel).
// Row blocks per column chunk of the wide product (a multiple of 8, about
// 4 SWEEP_RPW rows each on average).
int64_t wide_blocks(int64_t nf) {
  return ((nf + 4 * SWEEP_RPW - 1) / (4 * SWEEP_RPW) + 7) / 8 * 8;
}

// rblk (the wide product's row blocks, wide_blocks(nf) + 1 entries) is
// filled by row_blocks_kernel once per solve.
template <class TX>
int sweep_product(hipStream_t s, illum_ctx *ctx, const illum_sparse *F, int S, const TX *X,
                  const double *R, double *Y, const int *done, double *part, MixedIO m,
                  const int *rblk, const char *stage, bool timed) {
  StageTimer st(ctx, stage, timed);
  const int64_t nf = F->n;
  int nparts = 0;
  // Row blocks go to the XCDs as one contiguous range each (xc = 0), or
  // dealt in chunks of xc blocks.  The wide kernel's blocks are balanced by
  // weight, so ranges keep each XCD's gathers local without leaving one XCD
  // the crater floors' long rows.  Results do not depend on it: the wide
  // kernel tunes it, the others keep ranges.  ILLUM_RAD_XCD_CHUNK fixes it.
  const char *xenv = getenv("ILLUM_RAD_XCD_CHUNK");
  m.xc = xenv ? std::max(0, atoi(xenv)) : 0;
  if (S >= SWEEP_WIDE_MIN_S) {
    // PER: right-hand sides per lane.  The products of every PER are
    // identical (FMA accumulation in CSR order), so the autotuner times the
    // candidates (S % PER == 0, 64 PER <= S or PER = 1) on the solve's first
    // product when that product is idempotent (FP64: Y = R + F X; mixed ring
    // mode leaves B alone; mixed B mode adds to B, so it is not tuned).
    // Without it: the widest whose gathered chunk (nf x 64 PER values of X)
    // stays within 128 MB (ILLUM_SPMM_CHUNK_BYTES), so it is served from the
    // 256 MB memory-side cache rather than HBM.
    const bool idem = !(is_mixed<TX>() && m.B);
    const char *env = getenv("ILLUM_SPMM_CHUNK_BYTES");
    const double budget = env ? atof(env) : 128e6;
    int per = 8;
    while (per > 1 && (S % per != 0 || 64 * per > S ||
                       (double)nf * 64 * per * sizeof(TX) > budget))
      per >>= 1;
    const int nblk = (int)wide_blocks(nf);
    // Buffer gathers address X by 32-bit byte offsets: X must span < 2 GiB
    // (81,920 facets x 512 right-hand sides in FP64: 335 MB).
    const uint64_t xbytes = (uint64_t)nf * S * sizeof(TX);
    const bool buf = xbytes <= 0x7fffffffull;
    auto go = [&](int P) {
      const int nch = (S + 64 * P - 1) / (64 * P);
      const unsigned grid = (unsigned)((int64_t)nblk * nch);
#define ILLUM_SWEEP_WIDE(Q, B)                                                               \
  sweep_wide_kernel<TX, Q, B><<<grid, 256, 0, s>>>(nf, S, nblk, rblk, F->d_rowptr,            \
                                                   F->d_colidx, F->d_rval, X,                 \
                                                   (unsigned)xbytes, R, Y, done, part, m)
#define ILLUM_SWEEP_WIDE_P(Q)                                                                \
  do {                                                                                       \
    constexpr int xb = (int)sizeof(TX) * Q;                                                  \
    if constexpr (xb == 4 || xb == 8 || xb == 16) {                                          \
      if (buf) ILLUM_SWEEP_WIDE(Q, true); else ILLUM_SWEEP_WIDE(Q, false);                   \
    } else {                                                                                 \
      ILLUM_SWEEP_WIDE(Q, false);                                                            \
    }                                                                                        \
  } while (0)
      if (P == 8) ILLUM_SWEEP_WIDE_P(8);
      else if (P == 4) ILLUM_SWEEP_WIDE_P(4);
      else if (P == 2) ILLUM_SWEEP_WIDE_P(2);
      else ILLUM_SWEEP_WIDE_P(1);
#undef ILLUM_SWEEP_WIDE_P
#undef ILLUM_SWEEP_WIDE
      ILLUM_LAUNCH_CHECK();
    };
    std::vector<int> cands;
    for (int P : {8, 4, 2, 1})
      if (S % P == 0 && (64 * P <= S || P == 1)) cands.push_back(P);
    if (!env && idem && cands.size() > 1) {
      per = autotune(ctx, std::string("spmm_per[") + (is_mixed<TX>() ? "fp16" : "fp64") +
                              ",nf=" + std::to_string(nf) + ",S=" + std::to_string(S) + "]",
                     cands, per, go, timed);
    }
    if (!xenv && idem) {
      m.xc = autotune(ctx, std::string("spmm_xcd[") + (is_mixed<TX>() ? "fp16" : "fp64") +
                               ",nf=" + std::to_string(nf) + ",S=" + std::to_string(S) + "]",
                      {0, 1, 16}, 0, [&](int xc) { m.xc = xc; go(per); }, timed);
    }
    go(per);
    nparts = nblk;
  } else {
    const unsigned nb = (unsigned)((nf + SPMM_NNZ_ROWS - 1) / SPMM_NNZ_ROWS);
    sweep_nnz_kernel<TX><<<nb, 256, 0, s>>>(nf, S, F->d_rowptr, F->d_colidx, F->d_rval, X, R, Y,
                                            done, part, m);
    nparts = (int)nb;
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

// Corrections kept in a ring (S >= 128) and the sweeps between two folds.
constexpr int RAD_RING = 8;
constexpr int RAD_FOLD = RAD_RING - 1;

// Partial-sum blocks of the colnorm_partial_kernel pass (products that do
// not sum their own residual norms).
int colnorm_blocks(int64_t n) { return (int)std::min<int64_t>(1024, std::max<int64_t>(1, n / 16)); }

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte offsets of a solve's control workspace: control block, residual
// history, column norms of E and of the residual, scale exponents, remainder
// bounds, the wide product's row blocks and the partial column sums.
// (The wide product's partial sums take two sets of rows: the residuals'
// and, with FP32 remainders, the remainders'.)
struct CtlLayout {
  size_t hist, e2, col, cole, ex, tb, rblk, part2, part, bytes;
  CtlLayout(int64_t nf, int S, int max_iter) {
    const bool wide = S >= SWEEP_WIDE_MIN_S;
    const int64_t nblk = wide ? wide_blocks(nf) : 0;
    const int64_t gparts = std::max<int64_t>(
        {(int64_t)colnorm_blocks(nf * S), (nf + SPMM_NNZ_ROWS - 1) / SPMM_NNZ_ROWS, 2 * nblk});
    hist = 256;
    e2 = hist + align256(8 * ((size_t)max_iter + 1));
    col = e2 + align256(8 * (size_t)S);
    cole = col + align256(8 * (size_t)S);
    ex = cole + align256(8 * (size_t)S);
    tb = ex + align256(4 * ((size_t)max_iter + 2) * S);
    rblk = tb + align256(8 * (size_t)S);
    part2 = rblk + align256(4 * ((size_t)nblk + 1));
    part = part2 + align256(wide ? 8 * 2 * (size_t)WIDE_SUM_P * S : 0);
    bytes = part + 8 * (size_t)gparts * S;
  }
};

// FP32 remainders (ring mode) need a tolerance their rounding cannot reach:
// each stored remainder loses at most 2^-35 of the residual it came from, so
// their sum stays near 2^-35 ||E|| / (1 - rho(F)).  Tighter tolerances, and a
// solve whose rounding bound reaches the tolerance after all, use FP64
// residuals.
constexpr double RAD_REMAINDER_MIN_TOL = 1e-9;
// radiosity_solve flag (not part of the C ABI): FP64 residuals in ring mode.
constexpr int RAD_FP64_RESIDUALS = 1 << 30;

}  // namespace

// Scratch bytes of a solve (the pieces RadScratch hands out, each 256-byte
// aligned), for callers that own the buffer (a time loop).
size_t radiosity_scratch_bytes(int64_t nf, int S, int method, int max_iter) {
  const int64_t n = nf * S;
  const bool mixed = method == ILLUM_RAD_MIXED, ring = mixed && S >= RAD_RING_MIN_S;
  const int R = ring ? RAD_RING : 2;
  size_t b = align256(CtlLayout(nf, S, max_iter).bytes) + 2 * align256(8 * (size_t)n);
  if (mixed) b += align256(8 * (size_t)n) + align256(2 * (size_t)R * n);
  if (S > 1) b += align256(8 * (size_t)n);
  return b;
}

void radiosity_solve(illum_sparse *F, const double *E, int S, int method, double rel_tol,
                     int max_iter, double *B, double *resid, int *iters, int flags,
                     void *fixed_scratch, size_t fixed_cap, SolveStats *queue_stats) {
  illum_ctx *ctx = F->ctx;
  // fixed_scratch: the no-allocation variant (device E and B, bounded
  // max_iter, scratch carved from the fixed array, no profiling events)
  const bool noalloc = fixed_scratch != nullptr;
  // queue_stats: queue-only (a time loop's step, graph-capturable): exactly
  // max_iter products (those after convergence return at once), the output,
  // and the outcome folded into *queue_stats on the device; no host wait.
  const bool queue_only = queue_stats != nullptr;
  if (noalloc && (!(flags & ILLUM_IN_ON_DEVICE) || !(flags & ILLUM_OUT_ON_DEVICE)))
    throw Error(ILLUM_ERR_ARG, "the no-allocation solve takes E and B on the device");
  if (noalloc && !queue_only && max_iter > RAD_NOALLOC_MAX_ITER)
    throw Error(ILLUM_ERR_ARG, "the no-allocation solve takes max_iter <= " +
                                   std::to_string(RAD_NOALLOC_MAX_ITER));
  if (queue_only && !noalloc) throw Error(ILLUM_ERR_ARG, "a queued solve needs its own scratch");
  RadScratch scr{ctx, (unsigned char *)fixed_scratch, fixed_cap, 0};
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
  const bool wide = S >= SWEEP_WIDE_MIN_S;  // the wide product (row blocks, two-stage sums)
  // Mixed, S >= 128: the corrections stay in a ring and B is their sum
  // (folded every RAD_FOLD sweeps); otherwise each product adds its
  // correction to B ("B mode", two slots).
  const bool ring = mixed && S >= RAD_RING_MIN_S;
  const int R = ring ? RAD_RING : 2;

  // Scratch from the context's workspace (no allocation in a time loop):
  // control block, residual history, column norms, scale exponents and
  // partial sums.
  const int G = colnorm_blocks(n);
  const CtlLayout L(nf, S, max_iter);
  const size_t o_hist = L.hist;
  char *ws = (char *)scr.get(WS_RAD_CTL, L.bytes);
  SolveCtl *ctl = (SolveCtl *)ws;
  double *hist = (double *)(ws + L.hist), *e2 = (double *)(ws + L.e2),
         *col = (double *)(ws + L.col), *part = (double *)(ws + L.part),
         *part2 = (double *)(ws + L.part2);
  int *ex = mixed ? (int *)(ws + L.ex) : nullptr;
  // Ring mode keeps the residual as FP32 remainders q_k (host-driven solves
  // at a tolerance they can reach, ILLUM_RAD_REMAINDER=0/1 to force), with
  // the bound of their roundings in tb and their norms in cole.
  const char *renv = getenv("ILLUM_RAD_REMAINDER");
  const bool remainders =
      ring && !queue_only && !(flags & RAD_FP64_RESIDUALS) &&
      (renv ? atoi(renv) != 0 : rel_tol >= RAD_REMAINDER_MIN_TOL);
  double *tb = remainders ? (double *)(ws + L.tb) : nullptr;
  double *cole = remainders ? (double *)(ws + L.cole) : nullptr;
  int *rblk = wide ? (int *)(ws + L.rblk) : nullptr;
  const double sqrt_nf = std::sqrt((double)nf);
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
  double *X0 = (double *)scr.get(WS_RAD_X0, 8 * n);
  double *X1 = (double *)scr.get(WS_RAD_X1, 8 * n);
  // B mode accumulates B in place in the caller's B when it can (device,
  // S = 1, not overlapping E); ring mode folds into Bacc.
  double *Bacc = nullptr;
  _Float16 *Dring = nullptr;
  if (mixed) {
    Bacc = !ring && out_dev && S == 1 && !overlap(B, Ein, 8 * n)
               ? B : (double *)scr.get(WS_RAD_X2, 8 * n);
    Dring = (_Float16 *)scr.get(WS_RAD_D32, 2 * (size_t)R * n);
  }
  // Right-hand sides, row-major, for S > 1 (taken before any launch, so a
  // solve that does not fit the fixed scratch queues nothing).
  double *Erow = S > 1 ? (double *)scr.get(WS_RAD_E, 8 * n) : nullptr;

  // ||E_s||^2 from the column-major input (resetting the control block and,
  // mixed, writing the exponents of d_0 and d_1), then E row-major (with d_0 and B = 0 in the same pass).  With
  // S = 1 both layouts coincide and a device E is read in place (it is never
  // written).
  {
    const int P = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (nf + 4095) / 4096));
    const int64_t rows_per = (nf + P - 1) / P;
    colnorm_cm_partial_kernel<<<dim3((unsigned)S, (unsigned)P), 256, 0, s>>>(nf, S, rows_per,
                                                                            Ein, part);
    ILLUM_LAUNCH_CHECK();
    colnorm_final_kernel<<<(unsigned)((S + COLNORM_CW - 1) / COLNORM_CW), 256, 0, s>>>(
        P, S, COLNORM_CW, part, e2, nullptr, ctl, nullptr, 0.0, 0, ex, tb, nullptr, sqrt_nf);
    ILLUM_LAUNCH_CHECK();
  }
  if (wide) {
    const int nblk = (int)wide_blocks(nf);
    row_blocks_kernel<<<(unsigned)((nblk + 256) / 256), 256, 0, s>>>(nf, F->d_rowptr, nblk, rblk);
    ILLUM_LAUNCH_CHECK();
  }
  const double *Er = Ein;
  if (S > 1) {
    double *t = Erow;
    dim3 g((unsigned)((nf + 31) / 32), (unsigned)((S + 31) / 32));
    load_rhs_kernel<<<g, 256, 0, s>>>(nf, S, Ein, t, ex, mixed ? Dring : nullptr,
                                      mixed && !ring ? Bacc : nullptr);
    ILLUM_LAUNCH_CHECK();
    Er = t;
  } else if (mixed) {
    mixed_init_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, S, Er, ex, Dring, Bacc);
    ILLUM_LAUNCH_CHECK();
  }
  // The per-sweep column sums, COLNORM_CW columns per block (the wide
  // product's many partial rows first reduced to WIDE_SUM_P by
  // wide_colsum_kernel).
  const int cw = COLNORM_CW;
  const unsigned fb = (unsigned)((S + cw - 1) / cw);

  // Sweep k's buffers.  FP64 Jacobi, B_{k+1} = E + F B_k from B_0 = E:
  // cur_k = E, X1, X0, X1, ...; nxt_k = X1, X0, X1, ...  Mixed refinement,
  // r_0 = E: B mode r_k = E, X0, X1, X0, ...; r_{k+1} = X0, X1, X0, ...;
  // ring mode keeps the FP32 remainder q_k in X0 (k even) or X1 (k odd)
  // instead, q_0 unused (the first product forms eps_0 from E); d_k in slot
  // k % R, its exponents in ex[k].
  auto cur = [&](int k) -> const double * { return k == 0 ? Er : (k & 1 ? X1 : X0); };
  auto nxt = [&](int k) -> double * { return k & 1 ? X0 : X1; };
  auto rk = [&](int k) -> const double * { return k == 0 ? Er : (k & 1 ? X0 : X1); };
  auto rk1 = [&](int k) -> double * { return k & 1 ? X1 : X0; };
  auto qk = [&](int k) -> float * { return (float *)(k & 1 ? X1 : X0); };
  auto slot = [&](int j) { return Dring + (int64_t)(j % R) * n; };
  int folded = 0;  // ring mode: delta_j for j < min(folded, nprod) are in Bacc
  auto sweep = [&](int k) {
    if (ring && k > 0 && k % RAD_FOLD == 0) {
      fold_kernel<<<(unsigned)std::min<int64_t>((n + 255) / 256, 8192), 256, 0, s>>>(
          n, S, R, Dring, ex, ctl, k - RAD_FOLD, k, k == RAD_FOLD ? 1 : 0, Bacc);
      ILLUM_LAUNCH_CHECK();
      folded = k;
    }
    int nparts;
    const double *Y;
    const double *Xn = nullptr;  // subtracted in a separate residual norm (FP64: cur)
    if (!mixed) {
      // nxt = E + F cur; the residual of cur is nxt - cur
      nparts = sweep_product<double>(s, ctx, F, S, cur(k), Er, nxt(k), done, part, MixedIO{},
                                     rblk, "radiosity_fp64_sweep", !noalloc);
      Y = nxt(k);
      Xn = cur(k);
    } else {
      // r_{k+1} = (r_k - delta_k) + F delta_k, d_{k+1} = fp16(r_{k+1} 2^ex[k+1])
      MixedIO m{ex + (int64_t)k * S, ex + (int64_t)(k + 1) * S, slot(k + 1),
                ring ? nullptr : Bacc};
      if (remainders) {
        m.Qk = k == 0 ? nullptr : qk(k);
        m.Qn = qk(k + 1);
      }
      nparts = sweep_product<_Float16>(s, ctx, F, S, slot(k), rk(k),
                                       remainders ? nullptr : rk1(k), done, part, m, rblk,
                                       "radiosity_mixed_sweep", !noalloc);
      Y = rk1(k);
    }
    const double *psum = part;
    if (!nparts) {
      colnorm_partial_kernel<double><<<G, 256, 0, s>>>(nf, S, Y, Xn, part, done);
      ILLUM_LAUNCH_CHECK();
      nparts = G;
    } else if (wide) {
      wide_colsum_kernel<<<dim3((unsigned)((S + 63) / 64), WIDE_SUM_P, remainders ? 2 : 1), 256,
                           0, s>>>(nparts, S, part, part2, done);
      ILLUM_LAUNCH_CHECK();
      nparts = WIDE_SUM_P;
      psum = part2;
    }
    if (S == 1 && !tb && !cole) {
      colnorm_final1_kernel<<<1, 1024, 0, s>>>(nparts, psum, col, e2, ctl, hist, rel_tol, ex);
      ILLUM_LAUNCH_CHECK();
      return;
    }
    colnorm_final_kernel<<<fb, 256, 0, s>>>(nparts, S, cw, psum, col, e2, ctl, hist, rel_tol, 1,
                                            ex, tb, cole, sqrt_nf);
    ILLUM_LAUNCH_CHECK();
    if (fb > 1) {
      decide_kernel<<<1, 256, 0, s>>>(S, col, e2, ctl, hist, rel_tol, ex, tb, cole, sqrt_nf);
      ILLUM_LAUNCH_CHECK();
    }
  };

  // Sweeps are launched in batches (the first as long as the previous solve
  // of this matrix took); the control block and the residual history are
  // read back, in one copy, after each.
  // Control block and residual history read back to the host: on the stack
  // for the no-allocation variant (max_iter bounded), else on the heap.
  alignas(8) char hstack[256 + 8 * (RAD_NOALLOC_MAX_ITER + 1)];
  std::vector<char> hheap;
  if (!noalloc || queue_only) hheap.resize(queue_only ? 0 : o_hist + 8 * ((size_t)max_iter + 1));
  char *hbuf = noalloc ? hstack : hheap.data();
  SolveCtl h{};
  int launched = 0;
  int batch = F->last_nprod > 0 ? F->last_nprod : 4;
  // The output (column-major) is queued with each batch, ahead of the
  // readback; a batch that does not end the solve has its output rewritten
  // by the next one.
  DBuf<double> Bout;
  double *dst = B;
  if (!out_dev) { Bout.reset(n); dst = Bout.p; }
  auto output = [&]() {
    if (ring) {
      // B = sum of delta_j, j < nprod
      dim3 g((unsigned)((nf + 31) / 32), (unsigned)((S + 63) / 64));
      ring_out_kernel<<<g, 256, 0, s>>>(nf, S, R, Dring, ex, folded, ctl,
                                        folded > 0 ? Bacc : nullptr, dst);
      ILLUM_LAUNCH_CHECK();
    } else if (mixed) {
      if (S == 1) {
        if (Bacc != dst) ILLUM_HIP(hipMemcpyAsync(dst, Bacc, 8 * n, hipMemcpyDeviceToDevice, s));
      } else {
        transpose(s, nf, S, (const double *)Bacc, dst, false);
      }
    } else {
      // converged at sweep k -> B_k = cur(k); otherwise the newest iterate
      dim3 g((unsigned)((nf + 31) / 32), (unsigned)((S + 31) / 32));
      fp64_out_kernel<<<g, 256, 0, s>>>(nf, S, Er, X0, X1, ctl, dst);
      ILLUM_LAUNCH_CHECK();
    }
    if (!out_dev) ILLUM_HIP(hipMemcpyAsync(B, dst, 8 * n, hipMemcpyDeviceToHost, s));
  };
  if (queue_only) {
    for (int k = 0; k < max_iter; ++k) sweep(k);
    output();
    solve_stats_kernel<<<1, 1, 0, s>>>(ctl, hist, queue_stats);
    ILLUM_LAUNCH_CHECK();
    return;
  }
  // With S = 1 the sweeps read a device E in place: a B overlapping it is
  // written once the solve has ended.
  const bool early_out = !(S == 1 && overlap(dst, Ein, 8 * n));
  // The control block is device-written: its product count sizes the host
  // copies below, so it is checked against what was launched before use.
  auto check_ctl = [&]() {
    if (h.nprod < 0 || h.nprod > launched || launched > max_iter || (h.done != 0 && h.done != 1) ||
        (h.floor != 0 && (h.floor != 1 || !remainders)))
      throw Error(ILLUM_ERR_HIP, "radiosity solve: inconsistent device control block (" +
                                     std::to_string(h.nprod) + " products of " +
                                     std::to_string(launched) + " launched)");
  };
  for (;;) {
    const int nb = std::min(batch, max_iter - launched);
    for (int j = 0; j < nb; ++j) sweep(launched++);
    if (early_out) output();
    // launched <= max_iter: the copy fits hbuf (o_hist + 8 (max_iter + 1) bytes)
    ILLUM_HIP(hipMemcpyAsync(hbuf, ws, o_hist + 8 * (size_t)launched,
                             hipMemcpyDeviceToHost, s));
    sync(ctx);
    std::memcpy(&h, hbuf, sizeof(SolveCtl));
    check_ctl();
    if (h.done || launched >= max_iter) break;
    batch = std::min(2 * batch, 64);
  }
  // The FP32 remainders' rounding bound reached the tolerance: solve again
  // with FP64 residuals (whose output replaces this one's).
  if (h.floor)
    return radiosity_solve(F, E, S, method, rel_tol, max_iter, B, resid, iters,
                           flags | RAD_FP64_RESIDUALS, fixed_scratch, fixed_cap, queue_stats);
  if (!early_out) {
    output();
    if (!(in_dev && out_dev && (flags & ILLUM_ASYNC))) sync(ctx);
  }
  // 0 <= nprod <= launched <= max_iter: hh and the caller's resid (max_iter
  // + 1 entries) hold it
  const int nprod = h.nprod;
  F->last_nprod = nprod;
  if (resid && nprod > 0) std::memcpy(resid, hbuf + o_hist, 8 * (size_t)nprod);
  if (iters) *iters = nprod;
  double rel = 1.0;
  if (nprod > 0) std::memcpy(&rel, hbuf + o_hist + 8 * (size_t)(nprod - 1), 8);
  if (!(rel <= rel_tol))
    throw Error(ILLUM_ERR_NOT_CONVERGED,
                "radiosity solve did not reach the tolerance (relative residual " +
                    std::to_string(rel) + ")");
}

}  // namespace illum
