// Mesh-field transfer between shape models: k nearest facet centroids and the
// interpolation of per-facet fields (src/py/interpolate_mesh_field.py,
// src/py/build_mesh_transfer_tables.py, which query a scipy cKDTree).
//
// k-NN is exact brute force in FP64: every target scans every source point,
// with sources staged through LDS in tiles and each thread keeping its k best
// (squared distance, index) pairs sorted in registers.  Squared distances are
// ((dx dx + dy dy) + dz dz) with no fused operations; ties go to the smaller
// source index.  The k results come out nearest first with Euclidean
// distances sqrt(d2).
#include "hip_util.hpp"

#include <cmath>

namespace illum {

constexpr int KNN_BLOCK = 256;
constexpr int KNN_TILE = 1024;

template <int K>
__global__ __launch_bounds__(KNN_BLOCK) void knn_kernel(int64_t ns, const double *__restrict__ src,
                                                        int64_t nt, const double *__restrict__ tgt,
                                                        int k, int32_t *__restrict__ idx,
                                                        double *__restrict__ dist) {
  __shared__ double sx[KNN_TILE], sy[KNN_TILE], sz[KNN_TILE];
  const int64_t t = (int64_t)blockIdx.x * KNN_BLOCK + threadIdx.x;
  const bool act = t < nt;
  const double px = act ? tgt[3 * t] : 0.0, py = act ? tgt[3 * t + 1] : 0.0,
               pz = act ? tgt[3 * t + 2] : 0.0;
  double bd[K];
  int bi[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    bd[q] = __builtin_huge_val();
    bi[q] = 0x7fffffff;
  }
  for (int64_t s0 = 0; s0 < ns; s0 += KNN_TILE) {
    const int n = (int)(ns - s0 < KNN_TILE ? ns - s0 : KNN_TILE);
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += KNN_BLOCK) {
      sx[e] = src[3 * (s0 + e)];
      sy[e] = src[3 * (s0 + e) + 1];
      sz[e] = src[3 * (s0 + e) + 2];
    }
    __syncthreads();
    if (!act) continue;
    for (int e = 0; e < n; ++e) {
      const double dx = sx[e] - px, dy = sy[e] - py, dz = sz[e] - pz;
      const double d2 = (dx * dx + dy * dy) + dz * dz;
      const int i = (int)(s0 + e);
      // (d2, i) against the current k-th best; sources come in ascending
      // index, so an equal distance never displaces an earlier index
      if (d2 < bd[K - 1]) {
        double cd = d2;
        int ci = i;
#pragma unroll
        for (int q = 0; q < K; ++q) {
          if (cd < bd[q]) {
            const double td = bd[q];
            const int ti = bi[q];
            bd[q] = cd;
            bi[q] = ci;
            cd = td;
            ci = ti;
          }
        }
      }
    }
  }
  if (!act) return;
#pragma unroll
  for (int q = 0; q < K; ++q) {
    if (q < k) {
      idx[t * k + q] = bi[q];
      if (dist) dist[t * k + q] = sqrt(bd[q]);
    }
  }
}

// Field transfer for nlayers rows of a row-major [nlayers][n] field, the
// reference's three modes (interpolate_mesh_field.py):
//   nearest        X_tgt[l][i] = X_src[l][I[i][0]]
//   weighted       X_tgt[l][i] = sum_q Lam_q X_src[l][I[i][q]],
//                  Lam = D / sum(D) (weights grow with distance, as written)
//   area_weighted  X_tgt[l][i] = (sum_q Lam_q X_src[l][I_q] A_src[I_q]) / A_tgt[i]
// summed in q order.
__global__ __launch_bounds__(256) void transfer_kernel(int64_t nt, int nlayers, int64_t ns,
                                                       int mode, int k,
                                                       const int32_t *__restrict__ idx,
                                                       const double *__restrict__ dist,
                                                       const double *__restrict__ Xs,
                                                       const double *__restrict__ As,
                                                       const double *__restrict__ At,
                                                       double *__restrict__ Xt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int32_t *I = idx + t * k;
  const double *D = dist + t * k;
  if (mode == 0) {
    for (int l = 0; l < nlayers; ++l) Xt[(int64_t)l * nt + t] = Xs[(int64_t)l * ns + I[0]];
    return;
  }
  double sum = 0.0;
  for (int q = 0; q < k; ++q) sum += D[q];
  for (int l = 0; l < nlayers; ++l) {
    double acc = 0.0;
    for (int q = 0; q < k; ++q) {
      const double lam = D[q] / sum;
      const double x = Xs[(int64_t)l * ns + I[q]];
      acc += mode == 1 ? lam * x : lam * (x * As[I[q]]);
    }
    Xt[(int64_t)l * nt + t] = mode == 1 ? acc : acc / At[t];
  }
}

void knn(illum_ctx *ctx, const double *src, int64_t ns, const double *tgt, int64_t nt, int k,
         int32_t *idx, double *dist, int flags) {
  if (ns < 1 || nt < 0 || k < 1 || k > 32 || k > ns) throw Error(ILLUM_ERR_ARG, "bad k-NN sizes");
  if (ns >= ((int64_t)1 << 31)) throw Error(ILLUM_ERR_CAPACITY, "too many source points");
  if (!src || !tgt || !idx) throw Error(ILLUM_ERR_ARG, "null pointer");
  if (nt == 0) return;
  hipStream_t s = stream_of(ctx);
  const bool in_dev = flags & ILLUM_IN_ON_DEVICE, out_dev = flags & ILLUM_OUT_ON_DEVICE;
  DBuf<double> ds, dt, dd;
  DBuf<int32_t> di;
  const double *ps = src, *pt = tgt;
  if (!in_dev) {
    ds.reset(3 * ns);
    dt.reset(3 * nt);
    ILLUM_HIP(hipMemcpyAsync(ds.p, src, 24 * ns, hipMemcpyHostToDevice, s));
    ILLUM_HIP(hipMemcpyAsync(dt.p, tgt, 24 * nt, hipMemcpyHostToDevice, s));
    ps = ds.p;
    pt = dt.p;
  }
  int32_t *pi = idx;
  double *pd = dist;
  if (!out_dev) {
    di.reset(nt * k);
    pi = di.p;
    if (dist) {
      dd.reset(nt * k);
      pd = dd.p;
    }
  }
  {
    StageTimer st(ctx, "knn");
    const unsigned nb = (unsigned)((nt + KNN_BLOCK - 1) / KNN_BLOCK);
#define ILLUM_KNN(KK) knn_kernel<KK><<<nb, KNN_BLOCK, 0, s>>>(ns, ps, nt, pt, k, pi, pd)
    if (k <= 1) ILLUM_KNN(1);
    else if (k <= 2) ILLUM_KNN(2);
    else if (k <= 4) ILLUM_KNN(4);
    else if (k <= 8) ILLUM_KNN(8);
    else if (k <= 16) ILLUM_KNN(16);
    else ILLUM_KNN(32);
#undef ILLUM_KNN
    ILLUM_LAUNCH_CHECK();
  }
  if (!out_dev) {
    ILLUM_HIP(hipMemcpyAsync(idx, pi, 4 * nt * k, hipMemcpyDeviceToHost, s));
    if (dist) ILLUM_HIP(hipMemcpyAsync(dist, pd, 8 * nt * k, hipMemcpyDeviceToHost, s));
  }
  sync(ctx);
}

void mesh_transfer(illum_ctx *ctx, const double *src, int64_t ns, const double *tgt, int64_t nt,
                   const double *Xs, int nlayers, int mode, int k, const double *As,
                   const double *At, double *Xt) {
  if (mode < 0 || mode > 2) throw Error(ILLUM_ERR_ARG, "unknown transfer mode");
  if (nlayers < 1 || !Xs || !Xt) throw Error(ILLUM_ERR_ARG, "bad field");
  if (mode == 2 && (!As || !At)) throw Error(ILLUM_ERR_ARG, "area_weighted needs both areas");
  if (mode == 0) k = 1;
  if (nt == 0) return;
  hipStream_t s = stream_of(ctx);
  DBuf<double> ds(3 * ns), dtg(3 * nt), dd((size_t)nt * k), dXs((size_t)nlayers * ns),
      dXt((size_t)nlayers * nt), dAs(mode == 2 ? ns : 1), dAt(mode == 2 ? nt : 1);
  DBuf<int32_t> di((size_t)nt * k);
  ILLUM_HIP(hipMemcpyAsync(ds.p, src, 24 * ns, hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(dtg.p, tgt, 24 * nt, hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(dXs.p, Xs, 8 * (size_t)nlayers * ns, hipMemcpyHostToDevice, s));
  if (mode == 2) {
    ILLUM_HIP(hipMemcpyAsync(dAs.p, As, 8 * ns, hipMemcpyHostToDevice, s));
    ILLUM_HIP(hipMemcpyAsync(dAt.p, At, 8 * nt, hipMemcpyHostToDevice, s));
  }
  knn(ctx, ds.p, ns, dtg.p, nt, k, di.p, dd.p, ILLUM_IN_ON_DEVICE | ILLUM_OUT_ON_DEVICE);
  {
    StageTimer st(ctx, "transfer");
    transfer_kernel<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(nt, nlayers, ns, mode, k, di.p,
                                                                 dd.p, dXs.p, dAs.p, dAt.p, dXt.p);
    ILLUM_LAUNCH_CHECK();
  }
  ILLUM_HIP(hipMemcpyAsync(Xt, dXt.p, 8 * (size_t)nlayers * nt, hipMemcpyDeviceToHost, s));
  sync(ctx);
}

}  // namespace illum

using namespace illum;

extern "C" {

int illum_knn(illum_ctx *ctx, const double *src, int64_t ns, const double *tgt, int64_t nt, int k,
              int32_t *idx, double *dist, int flags) {
  return guarded_call([&] {
    if (!ctx) throw Error(ILLUM_ERR_ARG, "null context");
    ILLUM_HIP(hipSetDevice(ctx->device));
    knn(ctx, src, ns, tgt, nt, k, idx, dist, flags);
  });
}

int illum_mesh_transfer(illum_ctx *ctx, const double *src, int64_t ns, const double *tgt,
                        int64_t nt, const double *Xsrc, int nlayers, int mode, int k,
                        const double *area_src, const double *area_tgt, double *Xtgt) {
  return guarded_call([&] {
    if (!ctx) throw Error(ILLUM_ERR_ARG, "null context");
    ILLUM_HIP(hipSetDevice(ctx->device));
    mesh_transfer(ctx, src, ns, tgt, nt, Xsrc, nlayers, mode, k, area_src, area_tgt, Xtgt);
  });
}

}  // extern "C"
