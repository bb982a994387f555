// Form-factor assembly on the GPU (illum_context::compute_F, illum.cpp:31-237).
//
// Pipeline (all on the context stream):
//   1. ff_visibility: one workgroup per source facet j.  The O(N) facing cull
//      (illum.cpp:56-68) runs over all receivers with survivors compacted into
//      an LDS ray queue (wave ballot + popcount prefix), so the closest-hit
//      traversals (illum.cpp:72-94) run on full wavefronts.  Hit facets are
//      deduplicated in an LDS bitset of N bits (replacing the sorted insert of
//      illum.cpp:87-94).  The bitset row is written out raw (the visibility
//      set) and filtered by the form-factor value (zero drop, illum.cpp:125-127)
//      into the nonzero-pattern matrix M (row j = column j of F).
//   2. prefix sums of per-row counts give the CSC column pointers; ff_fill
//      walks each bit row in ascending order, so entries come out sorted by
//      (j, i) exactly as the reference sort (illum.cpp:184-191) leaves them.
//   3. a 512x512-bit tiled transpose of M yields the CSR pattern (row i =
//      receiver), filled the same way.  Everything is deterministic: no
//      floating-point atomics, no order-dependent accumulation.
#include "hip_util.hpp"
#include "ray.hpp"

#include <cfloat>
#include <cmath>
#include <algorithm>
#include <cstdlib>
#include <string>

namespace illum {

constexpr int FF_BLOCK = 256;
constexpr int FF_PER = 4;                          // candidates per thread per chunk
constexpr int FF_QCAP = FF_BLOCK * (FF_PER + 1);   // queue capacity (ints)
constexpr int FF_STACK = 8;  // LDS traversal stack entries per lane (deeper: threaded walk)
// First pass of each visibility ray prunes at this distance: the ray aims at a
// centroid lying on its target triangle, so the closest hit is almost always
// at t <= 1; a ray with no hit below the bound is traced again unbounded,
// which keeps the result identical to an unbounded closest-hit query.
constexpr float FF_TINIT = 1.0009765625f;
constexpr int FF_STAGE = 256;  // staged receivers per source row (CSC fill list)

// Form factor F(i, j) for receiver i and source j (illum.cpp:100-119), in FP64
// from the FP32 centroids/normals.  Returns 0 for r < DBL_EPSILON.
__device__ __forceinline__ double ff_value(float4 pi, float4 ni, float4 pj, float4 nj,
                                           double Aj) {
  const double px = (double)pj.x - (double)pi.x;
  const double py = (double)pj.y - (double)pi.y;
  const double pz = (double)pj.z - (double)pi.z;
  const double r = sqrt((px * px + py * py) + pz * pz);
  if (r < DBL_EPSILON) return 0.0;
  const double ux = px / r, uy = py / r, uz = pz / r;
  const double mij = fmax(0.0, ((double)ni.x * ux + (double)ni.y * uy) + (double)ni.z * uz);
  const double mji = fmax(0.0, -(((double)nj.x * ux + (double)nj.y * uy) + (double)nj.z * uz));
  return ((mij * mji) * Aj) / ((M_PI * r) * r);
}

// LAYOUT: 4 = near-first walk of the four-wide tree, 2 = of the two-wide
// tree, 0 = stackless threaded walk.  COUNT adds the work counters (pair, box
// and triangle tests) to stats[1..3]; stats[0] always receives the rays.
template <int LAYOUT, bool COUNT>
__global__ __launch_bounds__(FF_BLOCK) void ff_visibility_kernel(
    int nf, int nwords, const float4 *__restrict__ pr4, const float4 *__restrict__ n4,
    const float4 *__restrict__ nodes, const float4 *__restrict__ wide,
    const float4 *__restrict__ wide4, const float4 *__restrict__ tris, float offset,
    const double *__restrict__ area, const double *__restrict__ albedo_vec, double albedo,
    uint32_t *__restrict__ vis, uint32_t *__restrict__ mask, int32_t *__restrict__ counts,
    int *__restrict__ status, int nrows, int jstride, unsigned long long *__restrict__ stats,
    const float4 *__restrict__ clus, int nclus, int scap, const int32_t *__restrict__ body_of,
    const int32_t *__restrict__ body_off, const int32_t *__restrict__ roots,
    int32_t *__restrict__ col_counts, int32_t *__restrict__ stage) {
  extern __shared__ uint32_t smem[];
  uint32_t *bits = smem;
  int *queue = (int *)(smem + nwords);
  // surviving receiver clusters (16-bit ids: the LDS bitset bounds nf far
  // below 64 x 65536), then scap x FF_BLOCK stack slots (LAYOUT 2 or 4)
  uint16_t *clist = (uint16_t *)(queue + FF_QCAP);
  int *stack = (int *)(clist + ((nclus + 1) & ~1));
  __shared__ int s_qn, s_rays, s_nc;
  __shared__ int s_red[FF_BLOCK / 64], s_scan[FF_BLOCK / 64];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  // row = output row; source j = row * jstride (jstride 1 except when sampling)
  const int row = xcd_remap(blockIdx.x, nrows);
  const int j = row * jstride;

  for (int w = tid; w < nwords; w += FF_BLOCK) bits[w] = 0u;
  if (tid == 0) { s_qn = 0; s_rays = 0; s_nc = 0; }
  unsigned npair = 0, nbox = 0, ntri = 0;  // work counters of this thread
  // The source's body: receivers [lo, hi), its trees' roots.
  const int body = body_of ? body_of[j] : 0;
  const int lo = body_off[body], hi = body_off[body + 1];
  const int root_thr = roots[4 * body], root_wide = roots[4 * body + 1];
  const int root_wide4 = roots[4 * body + 2];
  const int clo = lo >> 6, chi = (hi + 63) >> 6;  // clusters overlapping [lo, hi)
  const float4 pj = pr4[j], nj = n4[j];
  const float qx = pj.x + offset * nj.x, qy = pj.y + offset * nj.y, qz = pj.z + offset * nj.z;
  __syncthreads();

  // Cluster pass: a cluster of 64 consecutive receivers whose bound proves
  // every member fails the facing test is dropped whole (the bound and its
  // rounding margin are documented with cull_clusters in mesh.cpp); the
  // others are listed for the per-pair test.  clus == nullptr keeps all.
  for (int c0 = clo; c0 < chi; c0 += FF_BLOCK) {
    const int c = c0 + tid;
    bool keep = false;
    if (c < chi) {
      keep = true;
      if (clus) {
        const float4 cc = clus[c];
        const float t = ((cc.x - pj.x) * nj.x + (cc.y - pj.y) * nj.y) + (cc.z - pj.z) * nj.z;
        keep = !(t <= -cc.w);
      }
    }
    const uint64_t m = __ballot(keep);
    if (m) {
      int slot = 0;
      if (lane == 0) slot = atomicAdd(&s_nc, __popcll(m));
      slot = __shfl(slot, 0);
      if (keep) clist[slot + __popcll(m & lanemask_lt())] = (uint16_t)c;
    }
  }
  __syncthreads();
  const int ncl = s_nc;

  // Per-pair facing test over the listed clusters, FF_PER * 4 clusters (one
  // per wave and k) per round; survivors go to the LDS ray queue.
  constexpr int CPR = FF_PER * (FF_BLOCK / 64);
  for (int g0 = 0; g0 < ncl; g0 += CPR) {
#pragma unroll
    for (int k = 0; k < FF_PER; ++k) {
      const int ci = g0 + k * (FF_BLOCK / 64) + wid;
      const int i = ci < ncl ? clist[ci] * 64 + lane : hi;
      bool s = false;
      if (i >= lo && i < hi && i != j) {
        ++npair;
        const float4 pi = pr4[i];
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float dot = (dx * nj.x + dy * nj.y) + dz * nj.z;
        s = !(dot <= -pi.w);
      }
      const uint64_t m = __ballot(s);
      if (m) {
        int slot = 0;
        if (lane == 0) {
          slot = atomicAdd(&s_qn, __popcll(m));
          atomicAdd(&s_rays, __popcll(m));
        }
        slot = __shfl(slot, 0);
        if (s) queue[slot + __popcll(m & lanemask_lt())] = i;
      }
    }
    __syncthreads();
    const bool last = g0 + CPR >= ncl;
    int qrem = s_qn;
    while (qrem >= FF_BLOCK || (last && qrem > 0)) {
      const int take = qrem < FF_BLOCK ? qrem : FF_BLOCK;
      if (LAYOUT == 3 || LAYOUT == 5) {
        // wavefront walk: the whole wavefront takes part, idle lanes masked;
        // a wavefront with no work left skips it
        const bool act = tid < take;
        if (__ballot(act)) {
          const int i = act ? queue[qrem - take + tid] : j;
          const float4 pi = pr4[i];
          const DRay r = make_ray(qx, qy, qz, pi.x - qx, pi.y - qy, pi.z - qz);
          float t;
          int h = LAYOUT == 3
                      ? trace_closest_packet(wide, tris, r, act, root_wide, FF_TINIT, &t, nbox, ntri)
                      : trace_closest_packet4(wide4, tris, r, act, root_wide4, FF_TINIT, &t, nbox,
                                              ntri);
          const bool again = act && h == -1;
          if (__ballot(again)) {
            const float inf = __builtin_huge_valf();
            const int h2 = LAYOUT == 3
                               ? trace_closest_packet(wide, tris, r, again, root_wide, inf, &t,
                                                      nbox, ntri)
                               : trace_closest_packet4(wide4, tris, r, again, root_wide4, inf, &t,
                                                       nbox, ntri);
            if (again) h = h2;
          }
          if (act && h >= 0 && h != j) atomicOr(&bits[h >> 5], 1u << (h & 31));
        }
      } else if (tid < take) {
        const int i = queue[qrem - take + tid];
        const float4 pi = pr4[i];
        const DRay r = make_ray(qx, qy, qz, pi.x - qx, pi.y - qy, pi.z - qz);
        float t;
        int h;
        if (LAYOUT == 4) {
          h = trace_closest_wide4(wide4, tris, r, root_wide4, FF_TINIT, stack + tid, FF_BLOCK,
                                  scap, &t, nbox, ntri);
          if (h == -1)
            h = trace_closest_wide4(wide4, tris, r, root_wide4, __builtin_huge_valf(),
                                    stack + tid, FF_BLOCK, scap, &t, nbox, ntri);
          // a walk deeper than the LDS stack: the stackless walk answers the
          // same unbounded query (identical result)
          if (h == -2) h = trace_closest(nodes, tris, r, root_thr, &t, nbox, ntri);
        } else if (LAYOUT == 2) {
          h = trace_closest_wide(wide, tris, r, root_wide, FF_TINIT, stack + tid, FF_BLOCK, scap,
                                 &t, nbox, ntri);
          if (h == -1)
            h = trace_closest_wide(wide, tris, r, root_wide, __builtin_huge_valf(), stack + tid,
                                   FF_BLOCK, scap, &t, nbox, ntri);
          if (h == -2) h = trace_closest(nodes, tris, r, root_thr, &t, nbox, ntri);
        } else {
          h = trace_closest(nodes, tris, r, root_thr, &t, nbox, ntri);
        }
        if (h >= 0 && h != j) atomicOr(&bits[h >> 5], 1u << (h & 31));
      }
      qrem -= take;
    }
    __syncthreads();
    if (tid == 0) s_qn = qrem;
    __syncthreads();
  }

  // Write the raw visibility row and its nonzero-form-factor filter.  A
  // row's set bits cluster in a few words, so their form factors are
  // evaluated one per thread from a list (in the now idle ray queue) rather
  // than by the words' owners.
  //   1. Each thread reads and clears its words, writes the raw row and lists
  //      the set bits; bits beyond the queue's capacity are evaluated by the
  //      owner at once, into its own (cleared) word.
  //   2. The listed receivers are evaluated, setting their filter bits by LDS
  //      atomics (no word is read as a visibility word any more).
  //   3. The filtered row is written.
  // The filter bits and counts do not depend on who evaluates what.
  const double Aj = area[j];
  bool nan = false;
  auto filter = [&](int i) -> bool {
    const double rho = albedo_vec ? albedo_vec[i] : albedo;
    const double v = ff_value(pr4[i], n4[i], pj, nj, Aj) * rho;
    if (v != v) nan = true;
    if (v == 0.0) return false;
    // receiver i's row count of the transposed matrix (the CSR row
    // lengths), so no pass over the transposed bits is needed for them
    if (col_counts) atomicAdd(&col_counts[i], 1);
    return true;
  };
  if (tid == 0) s_qn = 0;
  __syncthreads();
  for (int w = tid; w < nwords; w += FF_BLOCK) {
    const uint32_t word = bits[w];
    vis[(size_t)row * nwords + w] = word;
    uint32_t keep = 0u;
    if (word) {
      int slot = atomicAdd(&s_qn, __popc(word));
      uint32_t rest = word;
      while (rest) {
        const int b = __ffs(rest) - 1;
        rest &= rest - 1u;
        const int i = 32 * w + b;
        if (slot < FF_QCAP) queue[slot] = i;
        else if (filter(i)) keep |= 1u << b;
        ++slot;
      }
    }
    bits[w] = keep;
  }
  __syncthreads();
  const int nlist = min(s_qn, FF_QCAP);
  for (int e = tid; e < nlist; e += FF_BLOCK) {
    const int i = queue[e];
    if (filter(i)) atomicOr(&bits[i >> 5], 1u << (i & 31));
  }
  __syncthreads();
  int cnt = 0;
  for (int w = tid; w < nwords; w += FF_BLOCK) {
    const uint32_t keep = bits[w];
    mask[(size_t)row * nwords + w] = keep;
    cnt += __popc(keep);
  }
  if (nan) atomicOr(status, 1);
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
  if (lane == 0) s_red[wid] = cnt;
  __syncthreads();
  if (tid == 0) {
    int s = 0;
    for (int w = 0; w < FF_BLOCK / 64; ++w) s += s_red[w];
    counts[row] = s;
  }
  // 4. A row of at most FF_STAGE kept receivers is also listed, in ascending
  //    order, in its staging slot, so the CSC fill reads the short list
  //    rather than the bit row.  Threads own contiguous words; a block scan
  //    of their counts places each thread's receivers.
  if (stage) {
    int total = 0;
    for (int w = 0; w < FF_BLOCK / 64; ++w) total += s_red[w];
    if (total <= FF_STAGE) {  // uniform
      const int per = (nwords + FF_BLOCK - 1) / FF_BLOCK;
      const int w0 = min(nwords, tid * per), w1 = min(nwords, w0 + per);
      int c = 0;
      for (int w = w0; w < w1; ++w) c += __popc(bits[w]);
      int incl = c;
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
      }
      if (lane == 63) s_scan[wid] = incl;
      __syncthreads();
      int off = incl - c;
      for (int w = 0; w < wid; ++w) off += s_scan[w];
      int32_t *out = stage + (size_t)row * FF_STAGE;
      for (int w = w0; w < w1; ++w) {
        uint32_t rest = bits[w];
        while (rest) {
          const int b = __ffs(rest) - 1;
          rest &= rest - 1u;
          out[off++] = 32 * w + b;
        }
      }
    }
  }
  if (stats && tid == 0) atomicAdd(stats, (unsigned long long)s_rays);
  if (COUNT && stats) {
    // pair tests, box tests, triangle tests: wave sums, one atomic each.
    // Without COUNT the counters are dead and compile away.
    unsigned long long w3[3] = {npair, nbox, ntri};
    for (int q = 0; q < 3; ++q) {
      unsigned long long v = w3[q];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0 && v) atomicAdd(stats + 1 + q, v);
    }
  }
}

// Block-wide exclusive scan of one int per thread (blockDim == 256).
__device__ __forceinline__ int block_exscan256(int v, int *s_wave, int *total) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[wid] = x;
  __syncthreads();
  int off = 0, tot = 0;
  for (int w = 0; w < 4; ++w) {
    const int sw = s_wave[w];
    if (w < wid) off += sw;
    tot += sw;
  }
  __syncthreads();
  *total = tot;
  return off + x - v;
}

constexpr int FILL_MAXW = 16;

// Fills one CSC column (csr == 0: row of M is source j, bits are receivers i)
// or one CSR row (csr == 1: row of M^T is receiver i, bits are sources j).
// Rows of at most min_count entries are left to ff_fill_list_kernel.
__global__ __launch_bounds__(256) void ff_fill_kernel(
    int nrows, int nwords, const uint32_t *__restrict__ m, const int64_t *__restrict__ ptr,
    int csr, const float4 *__restrict__ pr4, const float4 *__restrict__ n4,
    const double *__restrict__ area, const double *__restrict__ albedo_vec, double albedo,
    int32_t *__restrict__ idx_out, double *__restrict__ val_out,
    const uint8_t *__restrict__ tile_nz, int64_t min_count) {
  __shared__ int s_wave[4];
  const int row = blockIdx.x;
  if (row >= nrows) return;
  if (ptr[row + 1] - ptr[row] <= min_count) return;  // filled from its staged list
  const int tid = threadIdx.x;
  const int per = (nwords + 255) / 256;
  const int w0 = tid * per, w1 = min(nwords, w0 + per);
  const uint32_t *mr = m + (size_t)row * nwords;
  // The thread's words stay in registers between the count and the listing
  // when there are at most FILL_MAXW of them (rows of up to 131,072 bits).
  const bool cached = per <= FILL_MAXW;
  // tile_nz (rows of the bit transpose): words of an all-zero 512 x 512 tile
  // were not written and read as zero
  const int ntile = nwords >> 4, tcol = row >> 9;
  auto word = [&](int w) -> uint32_t {
    return (!tile_nz || tile_nz[(w >> 4) * ntile + tcol]) ? mr[w] : 0u;
  };
  uint32_t wv[FILL_MAXW];
  int c = 0;
  if (cached) {
#pragma unroll
    for (int k = 0; k < FILL_MAXW; ++k) {
      wv[k] = w0 + k < w1 ? word(w0 + k) : 0u;
      c += __popc(wv[k]);
    }
  } else {
    for (int w = w0; w < w1; ++w) c += __popc(word(w));
  }
  int total;
  int off = block_exscan256(c, s_wave, &total);
  const int64_t base = ptr[row];
  // Phase 1: the row's set bits, in order, become its index list.
  int64_t pos = base + off;
  if (cached) {
#pragma unroll
    for (int k = 0; k < FILL_MAXW; ++k) {
      uint32_t rest = wv[k];
      while (rest) {
        const int b = __ffs(rest) - 1;
        rest &= rest - 1u;
        idx_out[pos++] = 32 * (w0 + k) + b;
      }
    }
  } else {
    for (int w = w0; w < w1; ++w) {
      uint32_t rest = word(w);
      while (rest) {
        const int b = __ffs(rest) - 1;
        rest &= rest - 1u;
        idx_out[pos++] = 32 * w + b;
      }
    }
  }
  // Phase 2: the form factors, one entry per thread (the set bits of a row
  // cluster in a few words, so phase 1 alone would leave them to one or two
  // threads).  The block reads back its own writes after the barrier, from
  // L2 (an agent-scope load passes the CU's L1).
  __syncthreads();
  const float4 prow = pr4[row], nrow = n4[row];
  for (int k = tid; k < total; k += 256) {
    const int col = __hip_atomic_load(&idx_out[base + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double v;
    if (csr) {  // row = receiver i, col = source j
      const double rho = albedo_vec ? albedo_vec[row] : albedo;
      v = ff_value(prow, nrow, pr4[col], n4[col], area[col]) * rho;
    } else {    // row = source j, col = receiver i
      const double rho = albedo_vec ? albedo_vec[col] : albedo;
      v = ff_value(pr4[col], n4[col], prow, nrow, area[row]) * rho;
    }
    val_out[base + k] = v;
  }
}

// CSC columns of at most FF_STAGE entries, from the receivers ff_visibility
// listed in ascending order in their staging slots: one wave per column,
// the same values as ff_fill_kernel's (ff_value on the same arguments).
__global__ __launch_bounds__(256) void ff_fill_list_kernel(
    int64_t nf, const int32_t *__restrict__ stage, const int64_t *__restrict__ colptr,
    const float4 *__restrict__ pr4, const float4 *__restrict__ n4,
    const double *__restrict__ area, const double *__restrict__ albedo_vec, double albedo,
    int32_t *__restrict__ rowidx, double *__restrict__ cval) {
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= nf) return;
  const int64_t base = colptr[j];
  const int64_t n = colptr[j + 1] - base;
  if (n > FF_STAGE) return;  // ff_fill_kernel reads this column's bit row
  const float4 pj = pr4[j], nj = n4[j];
  const double Aj = area[j];
  for (int e = threadIdx.x & 63; e < n; e += 64) {
    const int i = stage[(size_t)j * FF_STAGE + e];
    const double rho = albedo_vec ? albedo_vec[i] : albedo;
    rowidx[base + e] = i;
    cval[base + e] = ff_value(pr4[i], n4[i], pj, nj, Aj) * rho;
  }
}

// CSR(F) from CSC(F) when every row is short.  CSR row i lists the sources j
// of the CSC entries (i, j) in ascending j, with the same values (both fills
// evaluate ff_value on the same receiver / source arguments).
//
// csr_scatter_kernel: one wave per CSC column j; each entry (i, j, v) takes
// the next slot of row i through an atomic cursor, into the unsorted arrays
// tj / tv laid out by rowptr.
__global__ __launch_bounds__(256) void csr_scatter_kernel(
    int64_t nf, const int64_t *__restrict__ colptr, const int32_t *__restrict__ rowidx,
    const double *__restrict__ cval, const int64_t *__restrict__ rowptr,
    int32_t *__restrict__ cursor, int32_t *__restrict__ tj, double *__restrict__ tv) {
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= nf) return;
  for (int64_t k = colptr[j] + (threadIdx.x & 63); k < colptr[j + 1]; k += 64) {
    const int32_t i = rowidx[k];
    const int64_t at = rowptr[i] + atomicAdd(&cursor[i], 1);
    tj[at] = (int32_t)j;
    tv[at] = cval[k];
  }
}

// csr_sort_kernel: one wave per row of at most CSR_SORT_MAX entries, staged
// in LDS; an entry's place is the number of the row's entries with a smaller
// source index (all distinct), counted against LDS broadcasts.
constexpr int CSR_SORT_MAX = 1024;

__global__ __launch_bounds__(256) void csr_sort_kernel(int64_t nf,
                                                       const int64_t *__restrict__ rowptr,
                                                       const int32_t *__restrict__ tj,
                                                       const double *__restrict__ tv,
                                                       int32_t *__restrict__ colidx,
                                                       double *__restrict__ rval) {
  __shared__ int32_t sj[4][CSR_SORT_MAX];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + w;
  const bool valid = i < nf;
  const int64_t base = valid ? rowptr[i] : 0;
  const int n = valid ? (int)(rowptr[i + 1] - base) : 0;  // <= CSR_SORT_MAX (host check)
  for (int e = lane; e < n; e += 64) sj[w][e] = tj[base + e];
  __syncthreads();
  for (int e0 = 0; e0 < n; e0 += 64) {
    const int e = e0 + lane;
    const int32_t me = e < n ? sj[w][e] : 0;
    int rank = 0;
    for (int q = 0; q < n; ++q) rank += sj[w][q] < me ? 1 : 0;
    if (e < n) {
      colidx[base + rank] = me;
      rval[base + rank] = tv[base + e];
    }
  }
}

// Largest of n counts of a and of b, into out[0] and out[1] (zeroed by the
// caller): one atomic per block and array.
__global__ __launch_bounds__(256) void max_count_kernel(int64_t n, const int32_t *__restrict__ a,
                                                        const int32_t *__restrict__ b,
                                                        int32_t *__restrict__ out) {
  __shared__ int red[2][4];
  int ma = 0, mb = 0;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * 256) {
    ma = max(ma, a[k]);
    mb = max(mb, b[k]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    ma = max(ma, __shfl_xor(ma, o));
    mb = max(mb, __shfl_xor(mb, o));
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = ma;
    red[1][threadIdx.x >> 6] = mb;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const int m = max(max(red[threadIdx.x][0], red[threadIdx.x][1]),
                      max(red[threadIdx.x][2], red[threadIdx.x][3]));
    atomicMax(&out[threadIdx.x], m);
  }
}

// Transposes a (rows_p x 32*nwords)-bit matrix stored as rows of nwords words
// (rows_p == 32*nwords, both multiples of 512) in 512x512-bit tiles.  The
// tile is staged through LDS for coalesced 64-byte row segments; each 32x32
// sub-block (one word per lane, two sub-blocks per wavefront) is transposed in
// registers by the 5-stage block-swap network (Warren, Hacker's Delight,
// section 7-3) with cross-lane shuffles.
__global__ __launch_bounds__(256) void bit_transpose_kernel(const uint32_t *__restrict__ in,
                                                            uint32_t *__restrict__ out,
                                                            int nwords,
                                                            uint8_t *__restrict__ nz) {
  __shared__ uint32_t tile[512][17];
  const int tc = blockIdx.x, tr = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  uint32_t any = 0;
  for (int k = 0; k < 32; ++k) {
    const int idx = tid + 256 * k, r = idx >> 4, w = idx & 15;
    const uint32_t x = in[(size_t)(512 * tr + r) * nwords + 16 * tc + w];
    tile[r][w] = x;
    any |= x;
  }
  // An all-zero tile (most of them: a facet sees few others, mostly near it
  // in facet order) is only flagged; its output words are never read.
  const int nonzero = __syncthreads_or(any != 0);
  if (tid == 0) nz[tr * gridDim.x + tc] = nonzero ? 1 : 0;
  if (!nonzero) return;
  uint32_t res[32];
#pragma unroll
  for (int c = 0; c < 32; ++c) {
    const int combo = wv * 32 + c;  // (input word column w, pair of row blocks qp)
    const int w = combo >> 3, qp = combo & 7;
    const int q = 2 * qp + (lane >> 5);
    uint32_t x = tile[32 * q + (lane & 31)][w];
    uint32_t m = 0x0000FFFFu;
#pragma unroll
    for (int j = 16; j > 0; j >>= 1) {
      const uint32_t p = __shfl_xor(x, j);
      x = (lane & j) ? ((x & ~m) | ((p >> j) & m)) : ((x & m) | ((p & m) << j));
      m = m ^ (m << (j >> 1));
    }
    res[c] = x;  // lane l of the half holds column l of the sub-block
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 32; ++c) {
    const int combo = wv * 32 + c;
    const int w = combo >> 3, qp = combo & 7;
    const int q = 2 * qp + (lane >> 5);
    tile[32 * w + (lane & 31)][q] = res[c];
  }
  __syncthreads();
  for (int k = 0; k < 32; ++k) {
    const int idx = tid + 256 * k, o = idx >> 4, q = idx & 15;
    out[(size_t)(512 * tc + o) * nwords + 16 * tr + q] = tile[o][q];
  }
}


// Closest hit of arbitrary rays over every body's tree: each tree is walked
// unbounded and the (distance, index) minima are combined exactly.
__global__ __launch_bounds__(256) void closest_hit_kernel(
    int64_t nray, const float *__restrict__ orig, const float *__restrict__ dir,
    const float4 *__restrict__ nodes, const float4 *__restrict__ wide4,
    const float4 *__restrict__ tris, const int32_t *__restrict__ roots, int nbody,
    int32_t *__restrict__ hit, float *__restrict__ tout) {
  __shared__ int stack[FF_STACK * 256];
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nray) return;
  const DRay r = make_ray(orig[3 * k], orig[3 * k + 1], orig[3 * k + 2], dir[3 * k],
                          dir[3 * k + 1], dir[3 * k + 2]);
  unsigned nbox = 0, ntri = 0;
  int best = -1;
  float tbest = __builtin_huge_valf();
  for (int b = 0; b < nbody; ++b) {
    float t;
    int h = -2;
    if (wide4)
      h = trace_closest_wide4(wide4, tris, r, roots[4 * b + 2], __builtin_huge_valf(),
                              stack + threadIdx.x, 256, FF_STACK, &t, nbox, ntri);
    if (h == -2) h = trace_closest(nodes, tris, r, roots[4 * b], &t, nbox, ntri);  // deep walk
    if (h >= 0 && (t < tbest || (t == tbest && h < best))) {
      tbest = t;
      best = h;
    }
  }
  hit[k] = best;
  tout[k] = tbest;
}

// Traversal layout: the BVH's (two-wide near-first unless the autotuner chose
// four-wide), or ILLUM_FF_TRAVERSAL = "wide2", "wide4" or "threaded"
// (stackless).  The near-first walks keep an LDS stack of at most FF_STACK
// entries per lane; a walk that needs more is answered by the stackless
// threaded walk (same result).
// The wavefront walk needs the two-wide tree and depth + 1 <= 64 stack
// entries (one VGPR across the wavefront).
static bool packet_ok(const illum_bvh *bvh) { return bvh->d_wide && bvh->h.depth + 1 <= 64; }
static bool packet4_ok(const illum_bvh *bvh) {
  return bvh->d_wide4 && 3 * bvh->h.depth4 + 1 <= 64;
}

// Walk for a requested layout (2 two-wide, 3 wavefront packet over the
// two-wide tree, 4 four-wide, 5 wavefront packet over the four-wide tree, 0
// threaded), falling back to what the tree has.  ILLUM_FF_TRAVERSAL
// (threaded, wide2, packet, wide4, packet4) overrides.
static int ff_resolve(const illum_bvh *bvh, int want) {
  const char *e = getenv("ILLUM_FF_TRAVERSAL");
  const std::string ev = e ? e : "";
  if (ev == "threaded") want = 0;
  else if (ev == "wide2") want = 2;
  else if (ev == "packet") want = 3;
  else if (ev == "wide4") want = 4;
  else if (ev == "packet4") want = 5;
  if (want == 5 && packet4_ok(bvh)) return 5;
  if (want == 4 && bvh->d_wide4) return 4;
  if (want == 3 && packet_ok(bvh)) return 3;
  if (want == 0) return 0;
  return bvh->d_wide ? 2 : 0;
}

static int ff_layout(const illum_bvh *bvh) { return ff_resolve(bvh, bvh->ff_layout); }

// LDS stack entries a lane needs for a layout: one push per level (two-wide),
// up to three per level (four-wide), capped at FF_STACK.
static int ff_stack_slots(const illum_bvh *bvh, int layout) {
  if (layout == 2) return std::min(bvh->h.depth + 1, FF_STACK);
  if (layout == 4) return std::min(3 * bvh->h.depth4 + 1, FF_STACK);
  return 0;  // threaded walk; the wavefront walk keeps its stack in a VGPR
}

template <bool COUNT>
static decltype(&ff_visibility_kernel<4, COUNT>) ff_kernel(int layout) {
  return layout == 5   ? ff_visibility_kernel<5, COUNT>
         : layout == 4 ? ff_visibility_kernel<4, COUNT>
         : layout == 3 ? ff_visibility_kernel<3, COUNT>
         : layout == 2 ? ff_visibility_kernel<2, COUNT>
                       : ff_visibility_kernel<0, COUNT>;
}

// Dynamic LDS of ff_visibility_kernel: bitset, ray queue, 16-bit cluster list
// (rounded to whole words), stack.
static size_t ff_lds_bytes(int nwords, int64_t nclus, int stack_slots) {
  return (size_t)(nwords + FF_QCAP) * 4 + (size_t)((nclus + 1) & ~1) * 2 +
         (size_t)stack_slots * FF_BLOCK * 4;
}

// The cluster pass of the facing cull is on unless ILLUM_FF_CLUSTER_CULL=0
// (results are identical either way; the switch exists for comparison).
static bool cluster_cull() {
  const char *e = getenv("ILLUM_FF_CLUSTER_CULL");
  return !(e && std::string(e) == "0");
}

// ---------------------------------------------------------------------------
void bvh_closest_hit(illum_bvh *bvh, const float *orig, const float *dir, int64_t nray,
                     int32_t *hit, float *t) {
  illum_ctx *ctx = bvh->mesh->ctx;
  if (nray <= 0) return;
  DBuf<float> o(3 * nray), d(3 * nray), tt(nray);
  DBuf<int32_t> h(nray);
  hipStream_t s = stream_of(ctx);
  ILLUM_HIP(hipMemcpyAsync(o.p, orig, 12 * nray, hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(d.p, dir, 12 * nray, hipMemcpyHostToDevice, s));
  const int blocks = (int)((nray + 255) / 256);
  closest_hit_kernel<<<blocks, 256, 0, s>>>(
      nray, o.p, d.p, (const float4 *)bvh->d_nodes,
      ff_layout(bvh) == 4 ? (const float4 *)bvh->d_wide4 : nullptr, (const float4 *)bvh->d_tris,
      bvh->d_roots, bvh->nbody, h.p, tt.p);
  ILLUM_LAUNCH_CHECK();
  ILLUM_HIP(hipMemcpyAsync(hit, h.p, 4 * nray, hipMemcpyDeviceToHost, s));
  ILLUM_HIP(hipMemcpyAsync(t, tt.p, 4 * nray, hipMemcpyDeviceToHost, s));
  sync(ctx);
}

illum_sparse *form_factors(illum_bvh *bvh, double albedo, const double *albedo_vec,
                           double offset) {
  check_bvh_bodies(bvh);
  illum_mesh *mesh = bvh->mesh;
  illum_ctx *ctx = mesh->ctx;
  hipStream_t s = stream_of(ctx);
  const int64_t nf = mesh->g.nf;
  const int nwords = (int)(((nf + 511) / 512) * 16);
  const int64_t rows_p = (int64_t)nwords * 32;
  const int layout = ff_layout(bvh);
  // traversal stack: one slot per tree level (a push happens at most once per level)
  const int stack_slots = ff_stack_slots(bvh, layout);
  const size_t lds = ff_lds_bytes(nwords, mesh->nclus, stack_slots);
  if (lds > 160 * 1024)
    throw Error(ILLUM_ERR_CAPACITY, "too many facets for the LDS visibility bitset (max ~1.2M)");
  const void *kfn = (const void *)ff_kernel<false>(layout);
  if (lds > 64 * 1024)
    ILLUM_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));

  std::unique_ptr<illum_sparse, void (*)(illum_sparse *)> sp(new illum_sparse(), destroy_sparse);
  sp->ctx = ctx;
  sp->n = nf;
  sp->vis_words = nwords;
  sp->d_vis = palloc<uint32_t>(ctx, (size_t)nf * nwords);

  // per-facet albedo and the NaN status word in the context's workspace
  double *d_alb = nullptr;
  if (albedo_vec) {
    d_alb = (double *)ctx_workspace(ctx, WS_FF_ALBEDO, 8 * nf);
    ILLUM_HIP(hipMemcpyAsync(d_alb, albedo_vec, 8 * nf, hipMemcpyHostToDevice, s));
  }
  // status word (NaN seen), the largest CSR row and CSC column lengths
  int *status = (int *)ctx_workspace(ctx, WS_FF_STATUS, 12);
  int32_t *maxrow = (int32_t *)(status + 1);
  // each source's kept receivers, listed when at most FF_STAGE
  int32_t *stage = (int32_t *)ctx_workspace(ctx, WS_FF_STAGE, (size_t)nf * FF_STAGE * 4);
  // M (and, for rows too long to sort, its transpose) live in the context's
  // workspace.  Rows >= nf of M (padding to whole 512-row tiles) must read as
  // zero for the transpose.
  const size_t mbytes = (size_t)rows_p * nwords * 4;
  uint32_t *M = (uint32_t *)ctx_workspace(ctx, WS_FF_MASK, mbytes);
  // counts: CSC column lengths (rows of M), then CSR row lengths (rows of M^T)
  int32_t *counts = (int32_t *)ctx_workspace(ctx, WS_FF_COUNTS, 8 * rows_p);
  int32_t *rcounts = counts + rows_p;
  ILLUM_HIP(hipMemsetAsync(rcounts, 0, 4 * nf, s));
  if (rows_p > nf)
    ILLUM_HIP(hipMemsetAsync(M + (size_t)nf * nwords, 0, (size_t)(rows_p - nf) * nwords * 4, s));
  ILLUM_HIP(hipMemsetAsync(status, 0, 12, s));
  sp->d_nrays = palloc<unsigned long long>(ctx, 1);
  ILLUM_HIP(hipMemsetAsync(sp->d_nrays, 0, 8, s));

  {
    StageTimer st(ctx, "ff_visibility");
    auto launch = ff_kernel<false>(layout);
    launch<<<(int)nf, FF_BLOCK, lds, s>>>(
        (int)nf, nwords, (const float4 *)mesh->d_pr4, (const float4 *)mesh->d_n4,
        (const float4 *)bvh->d_nodes, (const float4 *)bvh->d_wide, (const float4 *)bvh->d_wide4,
        (const float4 *)bvh->d_tris,
        (float)offset, mesh->d_area, d_alb, albedo, sp->d_vis, M, counts, status,
        (int)nf, 1, sp->d_nrays, cluster_cull() ? (const float4 *)mesh->d_clus : nullptr,
        (int)mesh->nclus, stack_slots,
        mesh->d_body, mesh->d_boff, bvh->d_roots, rcounts, stage);
    ILLUM_LAUNCH_CHECK();
  }
  // CSC
  sp->d_colptr = palloc<int64_t>(ctx, nf + 1);
  {
    StageTimer st(ctx, "ff_csc");
    scan_counts(ctx, counts, nf, sp->d_colptr);
  }
  max_count_kernel<<<(unsigned)std::min<int64_t>(64, (nf + 255) / 256), 256, 0, s>>>(
      nf, rcounts, counts, maxrow);
  ILLUM_LAUNCH_CHECK();
  int64_t nnz = 0;
  int hstatus[3] = {0, 0, 0};  // NaN, longest CSR row, longest CSC column
  ILLUM_HIP(hipMemcpyAsync(&nnz, sp->d_colptr + nf, 8, hipMemcpyDeviceToHost, s));
  ILLUM_HIP(hipMemcpyAsync(hstatus, status, 12, hipMemcpyDeviceToHost, s));
  sync(ctx);
  if (hstatus[0]) throw Error(ILLUM_ERR_NAN, "computed a NaN form factor");
  if (nnz >= ((int64_t)1 << 40)) throw Error(ILLUM_ERR_CAPACITY, "nnz out of range");
  sp->nnz = nnz;
  sp->has_csc = true;
  sp->d_rowidx = palloc<int32_t>(ctx, nnz);
  sp->d_cval = palloc<double>(ctx, nnz);
  {
    // columns of at most FF_STAGE entries from their staged lists, longer
    // ones from their bit rows
    StageTimer st(ctx, "ff_csc");
    ff_fill_list_kernel<<<(unsigned)((nf + 3) / 4), 256, 0, s>>>(
        nf, stage, sp->d_colptr, (const float4 *)mesh->d_pr4, (const float4 *)mesh->d_n4,
        mesh->d_area, d_alb, albedo, sp->d_rowidx, sp->d_cval);
    ILLUM_LAUNCH_CHECK();
    if (hstatus[2] > FF_STAGE) {
      ff_fill_kernel<<<(int)nf, 256, 0, s>>>((int)nf, nwords, M, sp->d_colptr, 0,
                                             (const float4 *)mesh->d_pr4,
                                             (const float4 *)mesh->d_n4, mesh->d_area, d_alb,
                                             albedo, sp->d_rowidx, sp->d_cval, nullptr, FF_STAGE);
      ILLUM_LAUNCH_CHECK();
    }
  }
  sp->d_rowptr = palloc<int64_t>(ctx, nf + 1);
  sp->d_colidx = palloc<int32_t>(ctx, nnz + CSR_PAD);
  sp->d_rval = palloc<double>(ctx, nnz + CSR_PAD);
  ILLUM_HIP(hipMemsetAsync(sp->d_colidx + nnz, 0, 4 * CSR_PAD, s));
  ILLUM_HIP(hipMemsetAsync(sp->d_rval + nnz, 0, 8 * CSR_PAD, s));
  if (hstatus[1] <= CSR_SORT_MAX && !getenv("ILLUM_FF_CSR_TRANSPOSE")) {
    // CSR by a sparse transpose of the CSC entries (rows short enough to sort
    // in LDS); the transposed bit matrix's workspace holds the unsorted rows.
    StageTimer st(ctx, "ff_csr");
    scan_counts(ctx, rcounts, nf, sp->d_rowptr);
    const size_t tjb = ((size_t)4 * nnz + 255) & ~(size_t)255;
    char *tmp = (char *)ctx_workspace(ctx, WS_FF_MASK_T, tjb + 8 * (size_t)std::max<int64_t>(1, nnz));
    int32_t *tj = (int32_t *)tmp;
    double *tv = (double *)(tmp + tjb);
    ILLUM_HIP(hipMemsetAsync(counts, 0, 4 * nf, s));  // the CSC counts are spent: cursors
    const unsigned g4 = (unsigned)((nf + 3) / 4);
    csr_scatter_kernel<<<g4, 256, 0, s>>>(nf, sp->d_colptr, sp->d_rowidx, sp->d_cval,
                                          sp->d_rowptr, counts, tj, tv);
    ILLUM_LAUNCH_CHECK();
    csr_sort_kernel<<<g4, 256, 0, s>>>(nf, sp->d_rowptr, tj, tv, sp->d_colidx, sp->d_rval);
    ILLUM_LAUNCH_CHECK();
  } else {
    // CSR via the bit transpose.
    StageTimer st(ctx, "ff_csr");
    uint32_t *MT = (uint32_t *)ctx_workspace(ctx, WS_FF_MASK_T, mbytes);
    dim3 g(nwords / 16, (unsigned)(rows_p / 512));
    uint8_t *tile_nz = (uint8_t *)ctx_workspace(ctx, WS_FF_TILES, (size_t)g.x * g.y);
    bit_transpose_kernel<<<g, 256, 0, s>>>(M, MT, nwords, tile_nz);
    ILLUM_LAUNCH_CHECK();
    scan_counts(ctx, rcounts, nf, sp->d_rowptr);
    ff_fill_kernel<<<(int)nf, 256, 0, s>>>((int)nf, nwords, MT, sp->d_rowptr, 1,
                                           (const float4 *)mesh->d_pr4,
                                           (const float4 *)mesh->d_n4, mesh->d_area, d_alb,
                                           albedo, sp->d_colidx, sp->d_rval, tile_nz, -1);
    ILLUM_LAUNCH_CHECK();
  }
  sync(ctx);
  return sp.release();
}

// Times ff_visibility on a strided sample of sources for one BVH (ms, from
// hipEvents on the context stream).  Used by the leaf-size autotuner.
double time_visibility_sample(illum_bvh *bvh, int max_sources, int layout_choice) {
  check_bvh_bodies(bvh);
  illum_mesh *mesh = bvh->mesh;
  illum_ctx *ctx = mesh->ctx;
  hipStream_t s = stream_of(ctx);
  const int64_t nf = mesh->g.nf;
  const int jstride = (int)std::max<int64_t>(1, nf / std::max(1, max_sources));
  const int nrows = (int)((nf + jstride - 1) / jstride);
  const int nwords = (int)(((nf + 511) / 512) * 16);
  const int layout = ff_resolve(bvh, layout_choice);
  const int stack_slots = ff_stack_slots(bvh, layout);
  const size_t lds = ff_lds_bytes(nwords, mesh->nclus, stack_slots);
  if (lds > 160 * 1024) throw Error(ILLUM_ERR_CAPACITY, "too many facets for the LDS bitset");
  const void *kfn = (const void *)ff_kernel<false>(layout);
  if (lds > 64 * 1024)
    ILLUM_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  DBuf<uint32_t> vis((size_t)nrows * nwords), M((size_t)nrows * nwords);
  DBuf<int32_t> counts(nrows);
  DBuf<int> status(1);
  hipEvent_t e0, e1;
  ILLUM_HIP(hipEventCreate(&e0));
  ILLUM_HIP(hipEventCreate(&e1));
  auto launch = ff_kernel<false>(layout);
  float best = 1e30f;
  for (int rep = 0; rep < 2; ++rep) {  // first launch also loads the code object
    ILLUM_HIP(hipEventRecord(e0, s));
    launch<<<nrows, FF_BLOCK, lds, s>>>(
        (int)nf, nwords, (const float4 *)mesh->d_pr4, (const float4 *)mesh->d_n4,
        (const float4 *)bvh->d_nodes, (const float4 *)bvh->d_wide, (const float4 *)bvh->d_wide4,
        (const float4 *)bvh->d_tris,
        1e-5f, mesh->d_area, nullptr, 1.0, vis.p, M.p, counts.p, status.p, nrows, jstride, nullptr,
        cluster_cull() ? (const float4 *)mesh->d_clus : nullptr, (int)mesh->nclus, stack_slots,
        mesh->d_body, mesh->d_boff, bvh->d_roots, nullptr, nullptr);
    ILLUM_LAUNCH_CHECK();
    ILLUM_HIP(hipEventRecord(e1, s));
    ILLUM_HIP(hipEventSynchronize(e1));
    float ms = 0;
    ILLUM_HIP(hipEventElapsedTime(&ms, e0, e1));
    best = std::min(best, ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

// Work of one full visibility pass (all sources), counted by the COUNT
// instantiation: out4 = {rays, pair tests, box tests, triangle tests}.  The
// assembly kernel itself carries no counters; this is a separate launch.
void form_factor_work(illum_bvh *bvh, double offset, int64_t *out4) {
  check_bvh_bodies(bvh);
  illum_mesh *mesh = bvh->mesh;
  illum_ctx *ctx = mesh->ctx;
  hipStream_t s = stream_of(ctx);
  const int64_t nf = mesh->g.nf;
  const int nwords = (int)(((nf + 511) / 512) * 16);
  const int layout = ff_layout(bvh);
  const int stack_slots = ff_stack_slots(bvh, layout);
  const size_t lds = ff_lds_bytes(nwords, mesh->nclus, stack_slots);
  if (lds > 160 * 1024) throw Error(ILLUM_ERR_CAPACITY, "too many facets for the LDS bitset");
  const void *kfn = (const void *)ff_kernel<true>(layout);
  if (lds > 64 * 1024)
    ILLUM_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const size_t mbytes = (size_t)nf * nwords * 4;
  uint32_t *vis = (uint32_t *)ctx_workspace(ctx, WS_FF_MASK_T, mbytes);
  uint32_t *M = (uint32_t *)ctx_workspace(ctx, WS_FF_MASK, mbytes);
  int32_t *counts = (int32_t *)ctx_workspace(ctx, WS_FF_COUNTS, 4 * nf);
  DBuf<int> status(1);
  DBuf<unsigned long long> stats(4);
  ILLUM_HIP(hipMemsetAsync(stats.p, 0, 32, s));
  auto launch = ff_kernel<true>(layout);
  launch<<<(int)nf, FF_BLOCK, lds, s>>>(
      (int)nf, nwords, (const float4 *)mesh->d_pr4, (const float4 *)mesh->d_n4,
      (const float4 *)bvh->d_nodes, (const float4 *)bvh->d_wide, (const float4 *)bvh->d_wide4,
        (const float4 *)bvh->d_tris,
      (float)offset, mesh->d_area, nullptr, 1.0, vis, M, counts, status.p, (int)nf, 1, stats.p,
      cluster_cull() ? (const float4 *)mesh->d_clus : nullptr, (int)mesh->nclus, stack_slots,
        mesh->d_body, mesh->d_boff, bvh->d_roots, nullptr, nullptr);
  ILLUM_LAUNCH_CHECK();
  unsigned long long h[4];
  ILLUM_HIP(hipMemcpyAsync(h, stats.p, 32, hipMemcpyDeviceToHost, s));
  sync(ctx);
  for (int k = 0; k < 4; ++k) out4[k] = (int64_t)h[k];
}

}  // namespace illum
