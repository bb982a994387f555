// Form-factor assembly on the GPU (illum_context::compute_F, illum.cpp:31-237).
//
// Pipeline (all on the context stream):
//   1. ff_visibility: one workgroup per source facet j.  The O(N) facing cull
//      (illum.cpp:56-68) runs over the receivers of j's body with survivors
//      compacted into an LDS ray queue (wave ballot + popcount prefix), so the
//      closest-hit traversals (illum.cpp:72-94) run on full wavefronts.  Hit
//      facets are deduplicated in an LDS bitset over j's body (replacing the
//      sorted insert of illum.cpp:87-94) and filtered by the form-factor value
//      (zero drop, illum.cpp:125-127).  The kept receivers are appended to a
//      pool in ascending order, one atomic reservation per source, so the
//      device memory of an assembly is O(nnz), never O(N^2).  On request the
//      raw visibility sets go to a second pool the same way.
//   2. prefix sums of the per-source counts give the CSC column pointers; each
//      column is copied from its pool segment, already sorted by (j, i) as the
//      reference sort leaves it (illum.cpp:184-191), and its FP64 values
//      computed.
//   3. CSR(F) by a sparse transpose: entries scatter to their rows through
//      atomic cursors, then each row is ordered, in LDS by rank when short and
//      through an LDS bitmap over the row's body when long.
// Everything is deterministic: the only atomics are integer counts, pool
// reservations and cursors, whose order the per-column copy and the row
// ordering remove.
#include "hip_util.hpp"
#include "ray.hpp"

#include <cfloat>
#include <cmath>
#include <algorithm>
#include <cstdlib>
#include <array>
#include <string>

namespace illum {

constexpr int FF_BLOCK = 256;
constexpr int FF_MAXPER = 4;      // clusters per wavefront per fill round, at most
constexpr int FF_MAXCHUNK = 1024;  // rays per walk round, at most
#ifndef FF_KCELL
#define FF_KCELL 8
#endif
// direction keys: 6 permutation classes x FF_KCELL^2 direction cells
constexpr int FF_NKEY = 6 * FF_KCELL * FF_KCELL;
// Workgroups per CU the kernel is compiled for (8 wavefronts per SIMD: at
// most 64 VGPRs); the LDS per workgroup is sized for it where the body allows.
constexpr int FF_WAVES_PER_SIMD = 8;
constexpr int FF_STACK = 8;  // LDS traversal stack entries per lane (deeper: threaded walk)
// First pass of each visibility ray prunes at this distance: the ray aims at a
// centroid lying on its target triangle, so the closest hit is almost always
// at t <= 1; a ray with no hit below the bound is traced again unbounded,
// which keeps the result identical to an unbounded closest-hit query.
constexpr float FF_TINIT = 1.0009765625f;

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

// Arguments of one visibility launch (passed by value).
struct FFArgs {
  int nf;
  int nwmax;    // LDS bitset words: the largest body's facets / 32
  int nclmax;   // LDS cluster-list entries: the largest body's clusters
  int nrows;    // sources of the launch: j = row * jstride, or j = rows[row]
  int jstride;
  const int32_t *rows;  // nullable: the launch's sources, listed
  int scap;     // LDS traversal stack entries per lane
  int qchunk;   // rays per walk round (a multiple of FF_BLOCK, <= FF_MAXCHUNK)
  int fper;     // clusters per wavefront per fill round (<= FF_MAXPER)
  int qcap;     // ray queue entries: qchunk + one fill round (fper * 256)
  int qsort;    // 1: each round's rays are ordered by direction key first
  int clcap;    // cluster-list entries (the clusters are passed in chunks of it)
  // Paged bitset (PAGED launches): page table entries (pages of 1,024
  // receivers of the largest body), pool pages; sources that touch more
  // pages than the pool holds are listed in over_rows (count in *over_n)
  // with nothing else written, for a launch with a larger pool.
  int npgmax, npool;
  int32_t *over_rows;
  int *over_n;
  int xchunk;   // sources per chunk dealt to the XCDs in turn (0: one range per XCD)
  float offset;
  double albedo;
  const float4 *pr4, *n4, *nodes, *wide, *wide4, *tris, *clus;
  const double *area, *albedo_vec;
  const int32_t *body_of, *body_off, *roots;
  int *status;                   // bit 0: a NaN form factor
  unsigned long long *stats;     // rays [, pair, box, triangle tests with COUNT]
  int32_t *row_counts;           // CSR row lengths (receiver i), nullable
  // Kept receivers of each source, ascending, in a pool: counts[row]
  // entries from pool[offs[row] - off_base] (written when they fit below
  // pool_cap).
  int32_t *counts;
  int64_t *offs;
  int32_t *pool;
  unsigned long long *pool_n;
  int64_t pool_cap;
  int64_t off_base;
  // The raw visibility sets (before the zero drop) in the same form;
  // vpool == nullptr: not kept.
  int32_t *vcounts;
  int64_t *voffs;
  int32_t *vpool;
  unsigned long long *vpool_n;
  int64_t vpool_cap;
};

// Appends the set bits of an nw-word LDS row to a pool in ascending order:
// word w is word_at(w), its bit b is facet base_at(w) + b, and the words'
// facets ascend with w.  Threads own contiguous words, a block scan of their
// popcounts places each thread's entries, one atomic reserves the block's
// segment.  *cnt_out / *off_out receive the count and the segment's start
// plus off_base; nothing is written past cap (the host then runs the sources
// that did not fit again, into a second pool).  Called by the whole block.
template <class WordAt, class BaseAt>
__device__ void append_bits_sorted(WordAt word_at, BaseAt base_at, int nw, int32_t *pool,
                                   unsigned long long *pool_n, int64_t cap, int64_t off_base,
                                   int32_t *cnt_out, int64_t *off_out, int *s_scan,
                                   long long *s_base) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int per = (nw + FF_BLOCK - 1) / FF_BLOCK;
  const int w0 = min(nw, tid * per), w1 = min(nw, w0 + per);
  int c = 0;
  for (int w = w0; w < w1; ++w) c += __popc(word_at(w));
  int incl = c;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  if (lane == 63) s_scan[wid] = incl;
  __syncthreads();
  int off = incl - c, total = 0;
  for (int w = 0; w < FF_BLOCK / 64; ++w) {
    const int sw = s_scan[w];
    if (w < wid) off += sw;
    total += sw;
  }
  if (tid == 0) {
    const long long base = total ? (long long)atomicAdd(pool_n, (unsigned long long)total) : 0;
    *s_base = base;
    *cnt_out = total;
    *off_out = base + off_base;
  }
  __syncthreads();
  const long long base = *s_base;
  if (base + total <= cap) {
    int32_t *out = pool + base + off;
    for (int w = w0; w < w1; ++w) {
      uint32_t rest = word_at(w);
      const int fb = base_at(w);
      while (rest) {
        const int b = __ffs(rest) - 1;
        rest &= rest - 1u;
        *out++ = fb + b;
      }
    }
  }
  __syncthreads();  // s_scan and s_base are reused by the next call
}

// Direction key of a ray for grouping rays into wavefronts: its permutation
// class (ray_perm: the dominant axis and its sign) and a 4 x 4 cell of the
// two other components over the dominant one, Morton-ordered.  Only the
// grouping depends on it; each ray's result does not.
__device__ __forceinline__ int ray_dir_key(float dx, float dy, float dz) {
  const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
  int kz = 0;
  float m = ax;
  if (ay > m) { kz = 1; m = ay; }
  if (az > m) kz = 2;
  const float dk = kz == 0 ? dx : (kz == 1 ? dy : dz);
  const float du = kz == 0 ? dy : (kz == 1 ? dz : dx);
  const float dv = kz == 0 ? dz : (kz == 1 ? dx : dy);
  const float im = __builtin_amdgcn_rcpf(fabsf(dk));
  // cells of u, v in [-1, 1]; NaN (a zero direction) lands in cell 0
  constexpr float half = 0.5f * FF_KCELL;
  const int iu = (int)fminf(fmaxf((du * im + 1.0f) * half, 0.0f), FF_KCELL - 1.0f);
  const int iv = (int)fminf(fmaxf((dv * im + 1.0f) * half, 0.0f), FF_KCELL - 1.0f);
  int cell = 0;
#pragma unroll
  for (int b = 0; (1 << b) < FF_KCELL; ++b)
    cell |= (((iu >> b) & 1) << (2 * b)) | (((iv >> b) & 1) << (2 * b + 1));
  return (2 * kz + (dk < 0.0f ? 1 : 0)) * (FF_KCELL * FF_KCELL) + cell;
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

// Block-wide counting sort of n <= FF_MAXCHUNK queue entries (receivers
// relative to the body start lo) by the direction key of the ray from
// (qx, qy, qz) to the receiver's centroid, from `in` to `out`; the order
// within a key is arbitrary (LDS atomics).  `hist` holds FF_NKEY ints.
// Called by the whole block; ends with a barrier.
__device__ void sort_rays_by_key(const int32_t *in, int n, int32_t *out, int *hist,
                                 int *s_wave, const float4 *__restrict__ pr4, int lo, float qx,
                                 float qy, float qz) {
  constexpr int PER = FF_MAXCHUNK / FF_BLOCK;
  static_assert(FF_NKEY <= 2 * FF_BLOCK, "two histogram buckets per thread");
  const int tid = threadIdx.x;
  for (int k = tid; k < FF_NKEY; k += FF_BLOCK) hist[k] = 0;
  __syncthreads();
  int val[PER], key[PER], rk[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    const int e = tid + m * FF_BLOCK;
    key[m] = -1;
    val[m] = 0;
    rk[m] = 0;
    if (e < n) {
      val[m] = in[e];
      const float4 p = pr4[lo + val[m]];
      key[m] = ray_dir_key(p.x - qx, p.y - qy, p.z - qz);
      rk[m] = atomicAdd(&hist[key[m]], 1);
    }
  }
  __syncthreads();
  const int b0 = 2 * tid, b1 = 2 * tid + 1;
  const int h0 = b0 < FF_NKEY ? hist[b0] : 0, h1 = b1 < FF_NKEY ? hist[b1] : 0;
  int total;
  const int off = block_exscan256(h0 + h1, s_wave, &total);
  if (b0 < FF_NKEY) hist[b0] = off;
  if (b1 < FF_NKEY) hist[b1] = off + h0;
  __syncthreads();
#pragma unroll
  for (int m = 0; m < PER; ++m)
    if (key[m] >= 0) out[hist[key[m]] + rk[m]] = val[m];
  __syncthreads();
}

// Work counters of one thread (compiled away without COUNT): lane-level box
// and triangle tests, and wavefront-level walks, node visits, triangle tests
// and walks run with a fixed permutation class (equal on every lane).
struct FFWork {
  unsigned npair = 0, nbox = 0, ntri = 0, wwalk = 0, wnode = 0, wtri = 0, wspec = 0;
};

template <int LAYOUT, int PERM>
__device__ __forceinline__ int packet_walk_p(const FFArgs &a, const DRay &r, bool act, int root2,
                                             int root4, float tinit, float *t, FFWork &w) {
  if (LAYOUT == 3)
    return trace_closest_packet<PERM>(a.wide, a.tris, r, act, root2, tinit, t, w.nbox, w.ntri,
                                      w.wnode, w.wtri);
  return trace_closest_packet4<PERM>(a.wide4, a.tris, r, act, root4, tinit, t, w.nbox, w.ntri,
                                     w.wnode, w.wtri);
}

// Wavefront walk (LAYOUT 3: two-wide tree, 5: four-wide).  With SPEC, a
// wavefront whose active rays all share one permutation class (ray_perm)
// runs the walk compiled for that class, whose triangle test selects its
// axes at compile time; results are those of the general walk.  Called by
// the whole wavefront with at least one active lane.
template <int LAYOUT, bool SPEC>
__device__ __forceinline__ int packet_walk(const FFArgs &a, const DRay &r, bool act, int root2,
                                           int root4, float tinit, float *t, FFWork &w) {
  ++w.wwalk;
  if (SPEC) {
    const uint64_t am = wave_ballot(act);
    const int perm = ray_perm(r);
    const int p0 = __builtin_amdgcn_readlane(perm, (int)__builtin_ctzll(am));
    if (wave_ballot(act && perm != p0) == 0) {
      ++w.wspec;
      switch (p0) {
        case 0: return packet_walk_p<LAYOUT, 0>(a, r, act, root2, root4, tinit, t, w);
        case 1: return packet_walk_p<LAYOUT, 1>(a, r, act, root2, root4, tinit, t, w);
        case 2: return packet_walk_p<LAYOUT, 2>(a, r, act, root2, root4, tinit, t, w);
        case 3: return packet_walk_p<LAYOUT, 3>(a, r, act, root2, root4, tinit, t, w);
        case 4: return packet_walk_p<LAYOUT, 4>(a, r, act, root2, root4, tinit, t, w);
        default: return packet_walk_p<LAYOUT, 5>(a, r, act, root2, root4, tinit, t, w);
      }
    }
  }
  return packet_walk_p<LAYOUT, -1>(a, r, act, root2, root4, tinit, t, w);
}

// LAYOUT: 4 = near-first walk of the four-wide tree, 2 = of the two-wide
// tree, 3 / 5 = wavefront packet walks over the two- / four-wide tree, 0 =
// stackless threaded walk.  COUNT adds the work counters (pair, box and
// triangle tests) to stats[1..3]; stats[0] always receives the rays.
// PAGED keeps the visibility set in pages of 1,024 receivers allocated on
// first hit instead of one bit per receiver of the body, so the LDS does not
// grow with the body: for bodies whose bitset would limit the workgroups per
// CU.
template <int LAYOUT, bool COUNT, bool PAGED>
__global__ __launch_bounds__(FF_BLOCK, FF_WAVES_PER_SIMD) void ff_visibility_kernel(FFArgs a) {
  extern __shared__ uint32_t smem[];
  // LDS: the visibility set -- a bitset over the body (nwmax words), or
  // (PAGED) the page table (npgmax entries, -1 = no page), each pool page's
  // index and page number when listed, and the pool pages (32 words each) --
  // then the ray queue, the sorted walk round and key histogram (qsort), the
  // cluster list (16-bit offsets from its chunk's first cluster) and
  // scap x FF_BLOCK stack slots (LAYOUT 2 or 4).
  uint32_t *bits = smem;
  int *ptab = (int *)smem;
  int *pidx = ptab + a.npgmax;
  int *ppage = pidx + a.npool;
  uint32_t *pw = (uint32_t *)(ppage + a.npool);
  int *queue = PAGED ? (int *)(pw + 32 * a.npool) : (int *)(smem + a.nwmax);
  int *sorted = queue + a.qcap;
  int *hist = sorted + (a.qsort ? a.qchunk : 0);
  uint16_t *clist = (uint16_t *)(hist + (a.qsort ? FF_NKEY : 0));
  int *stack = (int *)(clist + ((a.clcap + 1) & ~1));
  __shared__ int s_qn, s_rays, s_nc, s_np, s_over;
  __shared__ float4 s_pn[2];  // source centroid, normal
  __shared__ int4 s_dims;     // body start and end, first cluster, source
  __shared__ int s_scan[FF_BLOCK / 64];
  __shared__ long long s_base;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  // row = output row; source j = row * jstride (jstride 1 except when sampling)
  const int r0 = a.xchunk > 0 ? xcd_chunked(blockIdx.x, a.nrows, a.xchunk)
                              : xcd_remap(blockIdx.x, a.nrows);
  const int row = a.rows ? a.rows[r0] : r0;
  const int j = row * a.jstride;
  // The source's body: receivers [lo, hi), its trees' roots.
  const int body = a.body_of ? a.body_of[j] : 0;
  const int lo = a.body_off[body], hi = a.body_off[body + 1];
  const int nw = (hi - lo + 31) >> 5;
  const int npg = (hi - lo + 1023) >> 10;
  if (PAGED) {
    for (int p = tid; p < npg; p += FF_BLOCK) ptab[p] = -1;
    for (int q = tid; q < 32 * a.npool; q += FF_BLOCK) pw[q] = 0u;
  } else {
    for (int w = tid; w < nw; w += FF_BLOCK) bits[w] = 0u;
  }
  if (tid == 0) { s_qn = 0; s_rays = 0; s_nc = 0; s_np = 0; s_over = 0; }
  FFWork w;  // work counters of this thread
  const int root_thr = a.roots[4 * body], root_wide = a.roots[4 * body + 1];
  const int root_wide4 = a.roots[4 * body + 2];
  const int clo = lo >> 6, chi = (hi + 63) >> 6;  // clusters overlapping [lo, hi)
  const float4 pj = a.pr4[j], nj = a.n4[j];
  const float qx = pj.x + a.offset * nj.x, qy = pj.y + a.offset * nj.y,
              qz = pj.z + a.offset * nj.z;
  if (tid == 0) {
    s_pn[0] = pj;
    s_pn[1] = nj;
    s_dims = make_int4(lo, hi, clo, j);
  }
  __syncthreads();

  // Marks receiver lo + rel as hit on the lanes with `hit`; called by the
  // whole wavefront.  PAGED: a page is taken from the pool on its first hit,
  // by one lane for all lanes of the wavefront that hit it; an LDS
  // compare-and-swap settles races between the block's 4 wavefronts, and the
  // index drawn by a loser is not reused (at most 3 lost per page: a source
  // hitting up to npool / 4 pages always fits).  A source needing more pages
  // than the pool holds is flagged.  (Claiming the entry before drawing the
  // index loses none but made the walk loop 25 % slower.)
  auto set_hit = [&](bool hit, int rel) {
    if (PAGED) {
      const int p = rel >> 10;
      int pg = hit ? ptab[p] : 0;
      uint64_t m = wave_ballot(hit && pg < 0);
      while (m) {
        const int l = (int)__builtin_ctzll(m);
        const int pl = __builtin_amdgcn_readlane(p, l);
        int got = 0;
        if (lane == l) {
          const int mine = atomicAdd(&s_np, 1);
          const int want = mine < a.npool ? mine : a.npool;
          const int old = atomicCAS(&ptab[pl], -1, want);
          got = old == -1 ? want : old;
        }
        got = __builtin_amdgcn_readlane(got, l);
        const bool same = hit && pg < 0 && p == pl;
        if (same) pg = got;
        m &= ~wave_ballot(same);
      }
      if (hit) {
        if (pg >= a.npool) s_over = 1;
        else atomicOr(&pw[32 * pg + ((rel >> 5) & 31)], 1u << (rel & 31));
      }
    } else if (hit) {
      atomicOr(&bits[rel >> 5], 1u << (rel & 31));
    }
  };

  // Walk rounds of up to qchunk rays from the queue's end (all of them with
  // flush); each wavefront takes 64 consecutive entries at a time.
  auto walk_rounds = [&](bool flush) {
    int qrem = s_qn;
    while (qrem >= a.qchunk || (flush && qrem > 0)) {
      const int n = qrem < a.qchunk ? qrem : a.qchunk;
      const int *src = queue + (qrem - n);
      if (a.qsort) {
        sort_rays_by_key(src, n, sorted, hist, s_scan, a.pr4, lo, qx, qy, qz);
        src = sorted;
      }
      for (int w0 = wid * 64; w0 < n; w0 += FF_BLOCK) {
        const int e = w0 + lane;
        const bool act = e < n;
        const int i = act ? lo + src[e] : j;
        const float4 pi = a.pr4[i];
        const DRay r = make_ray(qx, qy, qz, pi.x - qx, pi.y - qy, pi.z - qz);
        float t;
        int h = -1;
        if (LAYOUT == 3 || LAYOUT == 5) {
          // wavefront walk: the whole wavefront takes part, idle lanes masked
          h = packet_walk<LAYOUT, true>(a, r, act, root_wide, root_wide4, FF_TINIT, &t, w);
          const bool again = act && h == -1;
          if (wave_ballot(again)) {
            const int h2 = packet_walk<LAYOUT, false>(a, r, again, root_wide, root_wide4,
                                                      __builtin_huge_valf(), &t, w);
            if (again) h = h2;
          }
        } else if (act) {
          if (LAYOUT == 4) {
            h = trace_closest_wide4(a.wide4, a.tris, r, root_wide4, FF_TINIT, stack + tid,
                                    FF_BLOCK, a.scap, &t, w.nbox, w.ntri);
            if (h == -1)
              h = trace_closest_wide4(a.wide4, a.tris, r, root_wide4, __builtin_huge_valf(),
                                      stack + tid, FF_BLOCK, a.scap, &t, w.nbox, w.ntri);
            // a walk deeper than the LDS stack: the stackless walk answers
            // the same unbounded query (identical result)
            if (h == -2) h = trace_closest(a.nodes, a.tris, r, root_thr, &t, w.nbox, w.ntri);
          } else if (LAYOUT == 2) {
            h = trace_closest_wide(a.wide, a.tris, r, root_wide, FF_TINIT, stack + tid, FF_BLOCK,
                                   a.scap, &t, w.nbox, w.ntri);
            if (h == -1)
              h = trace_closest_wide(a.wide, a.tris, r, root_wide, __builtin_huge_valf(),
                                     stack + tid, FF_BLOCK, a.scap, &t, w.nbox, w.ntri);
            if (h == -2) h = trace_closest(a.nodes, a.tris, r, root_thr, &t, w.nbox, w.ntri);
          } else {
            h = trace_closest(a.nodes, a.tris, r, root_thr, &t, w.nbox, w.ntri);
          }
        }
        // h lies in [lo, hi): the walk covers only the body's tree
        set_hit(act && h >= 0 && h != j, h - lo);
      }
      qrem -= n;
      __syncthreads();  // the round's entries are read before the next round or fill
    }
    if (tid == 0) s_qn = qrem;
    __syncthreads();
  };

  // The receiver clusters in chunks of clcap (the whole body when its list
  // fits).  Cluster pass: a cluster of 64 consecutive receivers whose bound
  // proves every member fails the facing test is dropped whole (the bound
  // and its rounding margin are documented with cull_clusters in mesh.cpp);
  // the others are listed for the per-pair test.  clus == nullptr keeps all.
  const int cpr = a.fper * (FF_BLOCK / 64);
  for (int cc0 = clo; cc0 < chi; cc0 += a.clcap) {
    const int cc1 = min(chi, cc0 + a.clcap);
    for (int c0 = cc0; c0 < cc1; c0 += FF_BLOCK) {
      const int c = c0 + tid;
      bool keep = false;
      if (c < cc1) {
        keep = true;
        if (a.clus) {
          const float4 cc = a.clus[c];
          const float t = ((cc.x - pj.x) * nj.x + (cc.y - pj.y) * nj.y) + (cc.z - pj.z) * nj.z;
          keep = !(t <= -cc.w);
        }
      }
      const uint64_t m = wave_ballot(keep);
      if (m) {
        int slot = 0;
        if (lane == 0) slot = atomicAdd(&s_nc, __popcll(m));
        slot = __builtin_amdgcn_readfirstlane(slot);
        if (keep) clist[slot + rank_in_mask(m)] = (uint16_t)(c - cc0);
      }
    }
    __syncthreads();
    const int ncl = __builtin_amdgcn_readfirstlane(s_nc);

    // Per-pair facing test over the listed clusters, 4 fper clusters (one
    // per wave and k) per round; survivors go to the LDS ray queue.
    for (int g0 = 0; g0 < ncl; g0 += cpr) {
      // The source's point and normal and the body's range, read from LDS
      // for the round: kept in scalar registers they are spilled across the
      // walks and reloaded for every candidate.
      const float4 vp = s_pn[0], vn = s_pn[1];
      const int4 vd = s_dims;
      const float vpx = vp.x, vpy = vp.y, vpz = vp.z, vnx = vn.x, vny = vn.y, vnz = vn.z;
      const int vlo = vd.x, vhi = vd.y, vj = vd.w;
#pragma unroll
      for (int k = 0; k < FF_MAXPER; ++k) {
        if (k >= a.fper) break;
        const int ci = g0 + k * (FF_BLOCK / 64) + wid;
        const int i = ci < ncl ? (cc0 + (int)clist[ci]) * 64 + lane : vhi;
        const bool valid = i >= vlo && i < vhi && i != vj;
        if (valid) ++w.npair;
        // Every lane evaluates the test (a lane without a candidate reads
        // its own source and compares -inf, which fails), so the survivors'
        // mask is one comparison.
        const float4 pi = a.pr4[valid ? i : vj];
        const float dx = pi.x - vpx, dy = pi.y - vpy, dz = pi.z - vpz;
        const float dot = valid ? (dx * vnx + dy * vny) + dz * vnz : -__builtin_huge_valf();
        const bool sv = !(dot <= -pi.w);
        const uint64_t m = wave_ballot(sv);
        if (m) {
          int slot = 0;
          if (lane == 0) {
            slot = atomicAdd(&s_qn, __popcll(m));
            atomicAdd(&s_rays, __popcll(m));
          }
          slot = __builtin_amdgcn_readfirstlane(slot);
          if (sv) queue[slot + rank_in_mask(m)] = i - vlo;
        }
      }
      __syncthreads();
      walk_rounds(false);
    }
    if (tid == 0) s_nc = 0;  // the next chunk's list
    __syncthreads();
  }
  walk_rounds(true);

  // PAGED: a source that ran out of pool pages writes no lists (zero counts,
  // where the launch has them); it is listed for a launch with a bitset over
  // its body.  Its rays and work are counted here.
  const bool over = PAGED && __builtin_amdgcn_readfirstlane(s_over);
  if (over) {
    if (tid == 0) {
      if (a.counts) a.counts[row] = 0;
      if (a.offs) a.offs[row] = 0;
      if (a.vcounts) a.vcounts[row] = 0;
      if (a.voffs) a.voffs[row] = 0;
      if (a.over_rows) a.over_rows[atomicAdd(a.over_n, 1)] = row;
    }
  } else {
    // PAGED: the pool pages in ascending page order (ppage, pidx), by a block
    // scan over the page table; word w of the set is word w & 31 of listed
    // page w >> 5.
    int nvw = nw;
    if (PAGED) {
      const int per = (npg + FF_BLOCK - 1) / FF_BLOCK;
      const int p0 = min(npg, tid * per), p1 = min(npg, p0 + per);
      int c = 0;
      for (int p = p0; p < p1; ++p) c += ptab[p] >= 0 ? 1 : 0;
      int total;
      int off = block_exscan256(c, s_scan, &total);
      for (int p = p0; p < p1; ++p)
        if (ptab[p] >= 0) {
          ppage[off] = p;
          pidx[off] = ptab[p];
          ++off;
        }
      __syncthreads();
      nvw = 32 * total;
    }
    auto word_ref = [&](int vw) -> uint32_t & {
      return PAGED ? pw[32 * pidx[vw >> 5] + (vw & 31)] : bits[vw];
    };
    auto word_at = [&](int vw) -> uint32_t { return word_ref(vw); };
    auto base_at = [&](int vw) -> int {
      return PAGED ? lo + (ppage[vw >> 5] << 10) + ((vw & 31) << 5) : lo + 32 * vw;
    };
    auto facet_word = [&](int rel) -> uint32_t & {
      return PAGED ? pw[32 * ptab[rel >> 10] + ((rel >> 5) & 31)] : bits[rel >> 5];
    };

    // The raw visibility set, on request.
    if (a.vpool)
      append_bits_sorted(word_at, base_at, nvw, a.vpool, a.vpool_n, a.vpool_cap, 0,
                         a.vcounts + row, a.voffs + row, s_scan, &s_base);

    // Nonzero-form-factor filter.  A row's set bits cluster in a few words, so
    // their form factors are evaluated one per thread from a list (in the now
    // idle ray queue) rather than by the words' owners.
    //   1. Each thread reads and clears its words and lists the set bits; bits
    //      beyond the queue's capacity are evaluated by the owner at once, into
    //      its own (cleared) word.
    //   2. The listed receivers are evaluated, setting their filter bits by LDS
    //      atomics (no word is read as a visibility word any more).
    // The filter bits and counts do not depend on who evaluates what.
    const double Aj = a.area[j];
    bool nan = false;
    auto filter = [&](int i) -> bool {
      const double rho = a.albedo_vec ? a.albedo_vec[i] : a.albedo;
      const double v = ff_value(a.pr4[i], a.n4[i], pj, nj, Aj) * rho;
      if (v != v) nan = true;
      if (v == 0.0) return false;
      // receiver i's row count of the transposed matrix (the CSR row
      // lengths), so no pass over the transpose is needed for them
      if (a.row_counts) atomicAdd(&a.row_counts[i], 1);
      return true;
    };
    if (tid == 0) s_qn = 0;
    __syncthreads();
    for (int vw = tid; vw < nvw; vw += FF_BLOCK) {
      const uint32_t word = word_ref(vw);
      uint32_t keep = 0u;
      if (word) {
        int slot = atomicAdd(&s_qn, __popc(word));
        uint32_t rest = word;
        const int fb = base_at(vw);
        while (rest) {
          const int b = __ffs(rest) - 1;
          rest &= rest - 1u;
          const int i = fb + b;
          if (slot < a.qcap) queue[slot] = i;
          else if (filter(i)) keep |= 1u << b;
          ++slot;
        }
      }
      word_ref(vw) = keep;
    }
    __syncthreads();
    const int nlist = min(s_qn, a.qcap);
    for (int e = tid; e < nlist; e += FF_BLOCK) {
      const int i = queue[e];
      if (filter(i)) atomicOr(&facet_word(i - lo), 1u << ((i - lo) & 31));
    }
    if (nan) atomicOr(a.status, 1);
    __syncthreads();
    // The kept receivers, ascending.
    if (a.pool)
      append_bits_sorted(word_at, base_at, nvw, a.pool, a.pool_n, a.pool_cap, a.off_base,
                         a.counts + row, a.offs + row, s_scan, &s_base);
  }
  if (a.stats && tid == 0) atomicAdd(a.stats, (unsigned long long)s_rays);
  if (COUNT && a.stats) {
    // pair tests, box tests, triangle tests: wave sums, one atomic each;
    // the wavefront-level counters are equal on every lane (lane 0's).
    // Without COUNT the counters are dead and compile away.
    unsigned long long w3[3] = {w.npair, w.nbox, w.ntri};
    for (int q = 0; q < 3; ++q) {
      unsigned long long v = w3[q];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0 && v) atomicAdd(a.stats + 1 + q, v);
    }
    const unsigned long long w4[4] = {w.wwalk, w.wnode, w.wtri, w.wspec};
    for (int q = 0; q < 4; ++q)
      if (lane == 0 && w4[q]) atomicAdd(a.stats + 4 + q, w4[q]);
  }
}

// CSC(F) from the kept-receiver pools: one wave per column (= source j)
// copies its ascending segment to rowidx and evaluates the FP64 values.
__global__ __launch_bounds__(256) void ff_fill_list_kernel(
    int64_t nf, const int32_t *__restrict__ pool, int64_t cap, const int32_t *__restrict__ pool2,
    const int64_t *__restrict__ offs, const int64_t *__restrict__ colptr,
    const float4 *__restrict__ pr4,
    const float4 *__restrict__ n4, const double *__restrict__ area,
    const double *__restrict__ albedo_vec, double albedo, int32_t *__restrict__ rowidx,
    double *__restrict__ cval) {
  const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= nf) return;
  const int64_t base = colptr[j];
  const int64_t n = colptr[j + 1] - base;
  // offsets at or past cap are the second pool's (sources run again)
  const int64_t o = offs[j];
  const int32_t *src = o < cap ? pool + o : pool2 + (o - cap);
  const float4 pj = pr4[j], nj = n4[j];
  const double Aj = area[j];
  for (int64_t e = threadIdx.x & 63; e < n; e += 64) {
    const int i = src[e];
    const double rho = albedo_vec ? albedo_vec[i] : albedo;
    rowidx[base + e] = i;
    cval[base + e] = ff_value(pr4[i], n4[i], pj, nj, Aj) * rho;
  }
}

// Index lists from pool segments into one array laid out by ptr (the raw
// visibility sets): one wave per row.
__global__ __launch_bounds__(256) void list_copy_kernel(int64_t n, const int32_t *__restrict__ pool,
                                                        const int64_t *__restrict__ offs,
                                                        const int64_t *__restrict__ ptr,
                                                        int32_t *__restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const int64_t base = ptr[r], len = ptr[r + 1] - base;
  const int32_t *src = pool + offs[r];
  for (int64_t e = threadIdx.x & 63; e < len; e += 64) out[base + e] = src[e];
}

// Sources whose kept receivers did not fit a pool of cap entries (segment
// end past cap): listed in rows[0..*nrows) in any order, their counts summed
// into *need.
__global__ __launch_bounds__(256) void ff_overflow_rows_kernel(
    int64_t nf, const int32_t *__restrict__ counts, const int64_t *__restrict__ offs, int64_t cap,
    int32_t *__restrict__ rows, int *__restrict__ nrows, unsigned long long *__restrict__ need) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool over = j < nf && counts[j] > 0 && offs[j] + counts[j] > cap;
  const uint64_t m = wave_ballot(over);
  if (!m) return;
  const int lane = threadIdx.x & 63;
  int slot = 0;
  if (lane == (int)__builtin_ctzll(m)) slot = atomicAdd(nrows, __popcll(m));
  slot = __shfl(slot, (int)__builtin_ctzll(m));
  if (over) {
    rows[slot + rank_in_mask(m)] = (int32_t)j;
    atomicAdd(need, (unsigned long long)counts[j]);
  }
}

// CSR(F) from CSC(F).  CSR row i lists the sources j of the CSC entries
// (i, j) in ascending j, with the same values.
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

// csr_sort_kernel: one wave per row of at most max_len <= CSR_SORT_MAX
// entries, staged in LDS; an entry's place is the number of the row's entries
// with a smaller source index (all distinct), counted against LDS broadcasts.
constexpr int CSR_SORT_MAX = 1024;

__global__ __launch_bounds__(256) void csr_sort_kernel(int64_t nf, int64_t max_len,
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
  const int64_t len = valid ? rowptr[i + 1] - base : 0;
  const int n = len <= max_len ? (int)len : 0;  // longer rows: csr_long_kernel
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

// CSR rows longer than min_len: one block per row.  The row's unsorted
// source indices set bits of an LDS bitmap over a window of the row's index
// range [min, max] (at most 32 * wwords sources a window; the windows are
// taken in order), which is listed in order; each value is evaluated again
// with ff_value on the same receiver and source as the CSC entry, so it is
// bit-identical.  Shorter rows return at once (csr_sort_kernel orders them).
__global__ __launch_bounds__(256) void csr_long_kernel(
    int64_t min_len, int wwords, const int64_t *__restrict__ rowptr,
    const int32_t *__restrict__ tj, const float4 *__restrict__ pr4,
    const float4 *__restrict__ n4, const double *__restrict__ area,
    const double *__restrict__ albedo_vec, double albedo, int32_t *__restrict__ colidx,
    double *__restrict__ rval) {
  extern __shared__ uint32_t bm[];
  __shared__ int s_wave[4];
  __shared__ int s_min[4], s_max[4];
  const int64_t i = blockIdx.x;
  const int64_t base = rowptr[i], len = rowptr[i + 1] - base;
  if (len <= min_len) return;  // uniform over the block
  const int tid = threadIdx.x, lane = tid & 63;
  // the row's index range
  int mn = 0x7fffffff, mx = -1;
  for (int64_t e = tid; e < len; e += 256) {
    const int q = tj[base + e];
    mn = min(mn, q);
    mx = max(mx, q);
  }
  for (int o = 32; o > 0; o >>= 1) {
    mn = min(mn, __shfl_xor(mn, o));
    mx = max(mx, __shfl_xor(mx, o));
  }
  if (lane == 0) {
    s_min[tid >> 6] = mn;
    s_max[tid >> 6] = mx;
  }
  __syncthreads();
  const int rmin = min(min(s_min[0], s_min[1]), min(s_min[2], s_min[3]));
  const int rmax = max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
  const float4 pi = pr4[i], ni = n4[i];
  const double rho = albedo_vec ? albedo_vec[i] : albedo;
  int64_t pos = base;
  const int64_t wbits = 32 * (int64_t)wwords;
  for (int64_t w0 = rmin; w0 <= rmax; w0 += wbits) {
    const int64_t span = rmax - w0 + 1 < wbits ? rmax - w0 + 1 : wbits;
    const int nw = (int)((span + 31) >> 5);
    for (int w = tid; w < nw; w += 256) bm[w] = 0u;
    __syncthreads();
    for (int64_t e = tid; e < len; e += 256) {
      const int64_t q = tj[base + e] - w0;
      if (q >= 0 && q < 32 * (int64_t)nw) atomicOr(&bm[q >> 5], 1u << (q & 31));
    }
    __syncthreads();
    const int per = (nw + 255) / 256;
    const int v0 = min(nw, tid * per), v1 = min(nw, v0 + per);
    int c = 0;
    for (int w = v0; w < v1; ++w) c += __popc(bm[w]);
    int total;
    const int off = block_exscan256(c, s_wave, &total);
    int64_t p = pos + off;
    for (int w = v0; w < v1; ++w) {
      uint32_t rest = bm[w];
      while (rest) {
        const int b = __ffs(rest) - 1;
        rest &= rest - 1u;
        const int j = (int)(w0 + 32 * w + b);
        colidx[p] = j;
        rval[p] = ff_value(pi, ni, pr4[j], n4[j], area[j]) * rho;
        ++p;
      }
    }
    pos += total;
    __syncthreads();  // the bitmap is cleared for the next window
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

template <bool COUNT, bool PAGED>
static decltype(&ff_visibility_kernel<4, COUNT, PAGED>) ff_kernel_t(int layout) {
  return layout == 5   ? ff_visibility_kernel<5, COUNT, PAGED>
         : layout == 4 ? ff_visibility_kernel<4, COUNT, PAGED>
         : layout == 3 ? ff_visibility_kernel<3, COUNT, PAGED>
         : layout == 2 ? ff_visibility_kernel<2, COUNT, PAGED>
                       : ff_visibility_kernel<0, COUNT, PAGED>;
}
template <bool COUNT>
static decltype(&ff_visibility_kernel<4, COUNT, false>) ff_kernel(int layout, bool paged) {
  return paged ? ff_kernel_t<COUNT, true>(layout) : ff_kernel_t<COUNT, false>(layout);
}

// Per-source LDS of a mesh's bodies: the largest body's bitset words, pages
// of 1,024 receivers and cluster count (clusters overlapping its facet
// range).
struct BodyDims {
  int nwmax = 0, npgmax = 0, nclmax = 0;
};
static BodyDims body_dims(const illum_mesh *mesh) {
  BodyDims d;
  const auto &off = mesh->g.body_off;
  for (size_t b = 0; b + 1 < off.size(); ++b) {
    const int64_t lo = off[b], hi = off[b + 1];
    d.nwmax = std::max<int>(d.nwmax, (int)((hi - lo + 31) / 32));
    d.npgmax = std::max<int>(d.npgmax, (int)((hi - lo + 1023) / 1024));
    d.nclmax = std::max<int>(d.nclmax, (int)((hi + 63) / 64 - lo / 64));
  }
  return d;
}

// ILLUM_FF_PAGED: -1 unset (the shape chooser and the overflow rate
// decide), 0 bitset only, 1 paged set whenever a launch can take one.
static int ff_paged_env() {
  const char *pg = getenv("ILLUM_FF_PAGED");
  return pg ? (atoi(pg) ? 1 : 0) : -1;
}

// Pool pages of a paged launch: 32 or 16 (4 or 2 KB of visibility words).
// A source touching more pages of receivers runs again with a bitset over
// its body.  On icosphere-ordered meshes of 81,920 and 327,680 facets no
// source needs more than 16 (8 leave a few hundred sources over).
constexpr int FF_NPOOL_MAX = 32, FF_NPOOL_MIN = 16;

// LDS shape of a launch: bitset or paged set (pool pages), ray queue (rays
// per walk round, clusters per wavefront per fill round, capacity: a walk
// round plus a fill round), whether each round is sorted by direction, and
// the cluster-list entries.
struct QShape {
  int chunk = FF_BLOCK, fper = 1, cap = 2 * FF_BLOCK, sort = 0;
  int paged = 0, npool = 0, clcap = 0;
};

// Dynamic LDS of ff_visibility_kernel: visibility set (bitset, or page
// table, two page lists and the pool), ray queue, sorted round and key
// histogram (when sorting), 16-bit cluster list (rounded to whole words),
// stack.
static size_t ff_lds_bytes(const BodyDims &d, int stack_slots, const QShape &q) {
  const size_t set = q.paged ? (size_t)(d.npgmax + 2 * q.npool + 32 * q.npool) * 4
                             : (size_t)d.nwmax * 4;
  return set + (size_t)q.cap * 4 + (q.sort ? (size_t)(q.chunk + FF_NKEY) * 4 : 0) +
         (size_t)((q.clcap + 1) & ~1) * 2 + (size_t)stack_slots * FF_BLOCK * 4;
}

// Workgroups of one CU that a launch's LDS allows (160 KB per CU, in 512-byte
// granules, with the kernel's static LDS).
static int ff_groups_per_cu(size_t dyn) {
  const size_t per = (dyn + 256 + 511) & ~(size_t)511;
  return (int)std::min<size_t>(FF_WAVES_PER_SIMD, 160 * 1024 / per);
}

// The LDS shape: among a bitset or a paged set (pool of 32 or 16 pages),
// cluster lists of the whole body or chunks of 1,024, walk rounds of 1024 /
// 512 / 256 rays and fill rounds of 4 / 2 / 1 clusters per wavefront, the
// one with the most workgroups per CU, then the longest walk round (rays of
// a longer round sort into more coherent wavefronts: 6.1 against 6.7 ms at
// 81,920 facets), the longest fill round, a bitset, the larger pool and the
// longest list.  Results do not depend on it.  ILLUM_FF_RAY_SORT=0 walks
// rays in queue order; ILLUM_FF_RAY_CHUNK (256 / 512 / 1024) fixes the round
// size; ILLUM_FF_PAGED=0 / 1 fixes the set and ILLUM_FF_NPOOL the pages of
// its pool.
//   full: the shape that runs sources out of pool pages again -- a bitset,
// or for a body whose bitset exceeds the LDS the paged set with the largest
// pool that fits (then the most workgroups, ...).
static QShape ff_qshape(const BodyDims &d, int stack_slots, bool full = false) {
  const char *e = getenv("ILLUM_FF_RAY_SORT");
  const int sort = !(e && std::string(e) == "0");
  int fixed = 0;
  if (const char *c = getenv("ILLUM_FF_RAY_CHUNK")) {
    const int v = atoi(c);
    fixed = v >= 1024 ? 1024 : v >= 512 ? 512 : 256;
  }
  const int paged_only = ff_paged_env();
  int npool_fixed = 0;
  if (!full)
    if (const char *np = getenv("ILLUM_FF_NPOOL")) npool_fixed = std::max(1, atoi(np));
  QShape best;
  bool found = false;
  std::array<int, 6> best_score{};
  auto consider = [&](int paged, int np) {
    for (int clcap : {d.nclmax, std::min(d.nclmax, 1024)}) {
      if (clcap > 65536) continue;  // 16-bit list offsets
      for (int chunk = FF_MAXCHUNK; chunk >= FF_BLOCK; chunk /= 2) {
        if (fixed && chunk != fixed) continue;
        for (int fper = FF_MAXPER; fper >= 1; fper /= 2) {
          QShape q;
          q.chunk = chunk;
          q.fper = fper;
          q.cap = chunk + fper * FF_BLOCK;
          q.sort = sort;
          q.paged = paged;
          q.npool = np;
          q.clcap = clcap;
          const size_t lds = ff_lds_bytes(d, stack_slots, q);
          if (lds > 160 * 1024) continue;
          const int g = ff_groups_per_cu(lds);
          const std::array<int, 6> score =
              full ? std::array<int, 6>{np, g, chunk, fper, clcap, 0}
                   : std::array<int, 6>{g, chunk, fper, !paged, np, clcap};
          if (!found || score > best_score) {
            best = q;
            best_score = score;
            found = true;
          }
        }
      }
    }
  };
  if (full || paged_only != 1) consider(0, 0);
  if (full) {
    // no bitset fits: the largest pool, from the body's page count down
    if (!found && paged_only != 0)
      for (int np = d.npgmax; np >= 1 && !found; np = np > FF_NPOOL_MAX ? np * 3 / 4 : 0)
        consider(1, np);
  } else if (paged_only != 0) {
    if (npool_fixed) consider(1, std::min(npool_fixed, d.npgmax));
    else
      for (int np : {FF_NPOOL_MAX, FF_NPOOL_MIN}) consider(1, std::min(np, d.npgmax));
  }
  if (!found)
    throw Error(ILLUM_ERR_CAPACITY, "a body has too many facets for the LDS visibility set");
  return best;
}

// Checks the LDS a launch needs and raises the kernel's dynamic-LDS limit.
static size_t ff_prepare(const illum_mesh *mesh, const void *kfn, int stack_slots,
                         const QShape &q) {
  const BodyDims d = body_dims(mesh);
  const size_t lds = ff_lds_bytes(d, stack_slots, q);
  if (lds > 160 * 1024 || q.clcap > 65536)
    throw Error(ILLUM_ERR_CAPACITY, "a body has too many facets for the LDS visibility set");
  if (lds > 64 * 1024)
    ILLUM_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return lds;
}

static FFArgs ff_args(const illum_bvh *bvh, int nrows, int jstride, int stack_slots,
                      const QShape &q, double offset, const double *albedo_vec, double albedo) {
  const illum_mesh *mesh = bvh->mesh;
  const BodyDims d = body_dims(mesh);
  FFArgs a{};
  a.nf = (int)mesh->g.nf;
  a.nwmax = d.nwmax;
  a.nclmax = d.nclmax;
  a.nrows = nrows;
  a.jstride = jstride;
  a.scap = stack_slots;
  a.qchunk = q.chunk;
  a.fper = q.fper;
  // Sources are dealt to the XCDs in chunks of 8: the rays per source vary by
  // orders of magnitude between regions, and one contiguous range per XCD
  // left the busiest XCD with 20 % more rays than the mean at 81,920 facets
  // (9.2 ms against 6.8 ms).  ILLUM_FF_XCD_CHUNK overrides (0: one range
  // per XCD; 1: plain order).
  a.xchunk = 8;
  if (const char *e = getenv("ILLUM_FF_XCD_CHUNK")) a.xchunk = std::max(0, atoi(e));
  a.qcap = q.cap;
  a.qsort = q.sort;
  a.clcap = q.clcap;
  a.npgmax = d.npgmax;
  a.npool = q.npool;
  a.offset = (float)offset;
  a.albedo = albedo;
  a.pr4 = (const float4 *)mesh->d_pr4;
  a.n4 = (const float4 *)mesh->d_n4;
  a.nodes = (const float4 *)bvh->d_nodes;
  a.wide = (const float4 *)bvh->d_wide;
  a.wide4 = (const float4 *)bvh->d_wide4;
  a.tris = (const float4 *)bvh->d_tris;
  a.area = mesh->d_area;
  a.albedo_vec = albedo_vec;
  a.body_of = mesh->d_body;
  a.body_off = mesh->d_boff;
  a.roots = bvh->d_roots;
  return a;
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

// Pool capacity in entries: what the context's workspace already holds, at
// least `want`.
static int64_t pool_entries(illum_ctx *ctx, int slot, int64_t want) {
  const int64_t have = (int64_t)(slot < (int)ctx->workspace.size() ? ctx->workspace[slot].second : 0) / 4;
  return std::max(have, want);
}

illum_sparse *form_factors(illum_bvh *bvh, double albedo, const double *albedo_vec,
                           double offset, int flags) {
  check_bvh_bodies(bvh);
  illum_mesh *mesh = bvh->mesh;
  illum_ctx *ctx = mesh->ctx;
  hipStream_t s = stream_of(ctx);
  const int64_t nf = mesh->g.nf;
  const int layout = ff_layout(bvh);
  // traversal stack: one slot per tree level (a push happens at most once per level)
  const int stack_slots = ff_stack_slots(bvh, layout);
  const BodyDims dims = body_dims(mesh);
  // The shape that runs sources again (into the second pool, or after
  // running out of pool pages): a bitset, or for a body whose bitset exceeds
  // the LDS a paged set with the largest pool that fits.
  const QShape qf = ff_qshape(dims, stack_slots, true);
  auto launch_f = ff_kernel<false>(layout, qf.paged);
  const size_t lds_f = ff_prepare(mesh, (const void *)launch_f, stack_slots, qf);
  // The main shape.  Where a bitset fits, a paged set is dropped for this
  // tree once more than 1/64 of a pass's sources ran out of pool pages
  // (meshes whose facet order scatters a source's visible receivers over
  // many pages), unless ILLUM_FF_PAGED=1.
  const bool paged_forced = ff_paged_env() == 1;
  QShape qs = ff_qshape(dims, stack_slots);
  if (qs.paged && !qf.paged && bvh->ff_paged_off && !paged_forced) qs = qf;
  auto launch = ff_kernel<false>(layout, qs.paged);
  size_t lds = ff_prepare(mesh, (const void *)launch, stack_slots, qs);
  auto paged_overflow = [&](int64_t nover, int64_t nrows) {
    if (nover * 64 > nrows && !qf.paged && !paged_forced) bvh->ff_paged_off = true;
  };
  auto rerun_overflow = [&](int n) {
    if (n > 0)
      throw Error(ILLUM_ERR_CAPACITY,
                  "a source's visible receivers span more pages of 1,024 facets than the "
                  "LDS holds");
  };
  const bool keep_vis = flags & ILLUM_FF_KEEP_VISIBILITY;

  std::unique_ptr<illum_sparse, void (*)(illum_sparse *)> sp(new illum_sparse(), destroy_sparse);
  sp->ctx = ctx;
  sp->n = nf;

  // per-facet albedo in the context's workspace
  double *d_alb = nullptr;
  if (albedo_vec) {
    d_alb = (double *)ctx_workspace(ctx, WS_FF_ALBEDO, 8 * nf);
    ILLUM_HIP(hipMemcpyAsync(d_alb, albedo_vec, 8 * nf, hipMemcpyHostToDevice, s));
  }
  // Control words: NaN status, the longest CSR row and CSC column, the two
  // pool counters.
  struct Ctl {
    int status, maxrow, maxcol, nover;
    unsigned long long pool_n, vpool_n, need, pool2_n;
    // sources out of pool pages (paged launches): main pass, reruns
    int nover_pg, nover_pg2;
  };
  Ctl *ctl = (Ctl *)ctx_workspace(ctx, WS_FF_STATUS, sizeof(Ctl));
  // Per-source counts and pool offsets, CSR row lengths: 24 bytes per facet
  // (28 with the raw visibility sets).
  char *cw = (char *)ctx_workspace(ctx, WS_FF_COUNTS, (size_t)nf * (keep_vis ? 40 : 24));
  int64_t *offs = (int64_t *)cw;
  int32_t *counts = (int32_t *)(cw + 8 * nf);
  int32_t *rcounts = (int32_t *)(cw + 12 * nf);
  int64_t *voffs = keep_vis ? (int64_t *)(cw + 16 * nf) : nullptr;
  int32_t *vcounts = keep_vis ? (int32_t *)(cw + 24 * nf) : nullptr;
  sp->d_nrays = palloc<unsigned long long>(ctx, 1);

  // Pool capacity: the last assembly's nonzeros on this tree plus 5 %, else
  // an estimate from a counted sample of every 32nd source (nothing written)
  // plus 25 %; at least what the context's workspace already holds.  Sources
  // whose lists still do not fit run again into a second pool sized exactly.
  int32_t *over_rows = (int32_t *)ctx_workspace(ctx, WS_FF_OVER, 4 * (size_t)nf);
  auto base_args = [&](int nrows, int jstride, const QShape &q) {
    FFArgs a = ff_args(bvh, nrows, jstride, stack_slots, q, offset, d_alb, albedo);
    a.status = &ctl->status;
    a.clus = cluster_cull() ? (const float4 *)mesh->d_clus : nullptr;
    a.counts = counts;
    a.offs = offs;
    a.over_rows = over_rows;
    a.over_n = &ctl->nover_pg;
    return a;
  };
  int64_t want = 0;
  if (bvh->ff_nnz_last > 0) {
    want = bvh->ff_nnz_last + bvh->ff_nnz_last / 20 + 4096;
  } else {
    const int js = nf >= 32 * 64 ? 32 : 1;
    const int nrows = (int)((nf + js - 1) / js);
    ILLUM_HIP(hipMemsetAsync(ctl, 0, sizeof(Ctl), s));
    FFArgs a = base_args(nrows, js, qs);
    a.pool = counts;  // capacity 0: only the counts are written
    a.pool_n = &ctl->pool_n;
    a.pool_cap = 0;
    launch<<<nrows, FF_BLOCK, lds, s>>>(a);
    ILLUM_LAUNCH_CHECK();
    Ctl h0{};
    ILLUM_HIP(hipMemcpyAsync(&h0, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
    sync(ctx);
    const double est = (double)h0.pool_n * (double)nf / (double)nrows;
    want = (int64_t)(est * 1.25) + 4096;
    if (qs.paged) {
      // sources out of pool pages count nothing in the sample: with many of
      // them the pass keeps a bitset, and the estimate is scaled up
      paged_overflow(h0.nover_pg, nrows);
      if (bvh->ff_paged_off && !qf.paged) {
        qs = qf;
        launch = launch_f;
        lds = lds_f;
      }
      const int64_t ok = std::max<int64_t>(1, nrows - h0.nover_pg);
      want = (int64_t)(est * 1.25 * (double)nrows / (double)ok) + 4096;
    }
  }
  int64_t cap = pool_entries(ctx, WS_FF_POOL, want);
  // ILLUM_FF_POOL_ENTRIES fixes the first pool's capacity (to exercise the
  // second pool; results are identical).
  if (const char *e = getenv("ILLUM_FF_POOL_ENTRIES")) cap = std::max<int64_t>(0, atoll(e));
  int64_t vcap = keep_vis ? pool_entries(ctx, WS_FF_VPOOL, 2 * want) : 0;
  int64_t nnz = 0, cap2 = 0;
  Ctl h{};
  sp->d_colptr = palloc<int64_t>(ctx, nf + 1);
  int32_t *pool = nullptr, *vpool = nullptr, *pool2 = nullptr;
  for (int pass = 0;; ++pass) {
    pool = (int32_t *)ctx_workspace(ctx, WS_FF_POOL, 4 * (size_t)std::max<int64_t>(1, cap));
    if (keep_vis) vpool = (int32_t *)ctx_workspace(ctx, WS_FF_VPOOL, 4 * (size_t)vcap);
    ILLUM_HIP(hipMemsetAsync(rcounts, 0, 4 * nf, s));
    ILLUM_HIP(hipMemsetAsync(ctl, 0, sizeof(Ctl), s));
    ILLUM_HIP(hipMemsetAsync(sp->d_nrays, 0, 8, s));
    FFArgs a = base_args((int)nf, 1, qs);
    a.stats = sp->d_nrays;
    a.row_counts = rcounts;
    a.pool = pool;
    a.pool_n = &ctl->pool_n;
    a.pool_cap = cap;
    a.vcounts = vcounts;
    a.voffs = voffs;
    a.vpool = vpool;
    a.vpool_n = &ctl->vpool_n;
    a.vpool_cap = vcap;
    {
      StageTimer st(ctx, "ff_visibility");
      launch<<<(int)nf, FF_BLOCK, lds, s>>>(a);
      ILLUM_LAUNCH_CHECK();
    }
    // Counts scanned to column offsets, the longest row and column, the
    // sources whose lists did not fit the pool (none when it sufficed), and
    // the control words, in one host wait.
    int32_t *orows = (int32_t *)ctx_workspace(ctx, WS_FF_ROWS, 4 * (size_t)nf);
    auto scan_and_read = [&]() {
      {
        StageTimer st(ctx, "ff_csc");
        scan_counts(ctx, counts, nf, sp->d_colptr);
      }
      max_count_kernel<<<(unsigned)std::min<int64_t>(64, (nf + 255) / 256), 256, 0, s>>>(
          nf, rcounts, counts, &ctl->maxrow);
      ILLUM_LAUNCH_CHECK();
      ff_overflow_rows_kernel<<<(unsigned)((nf + 255) / 256), 256, 0, s>>>(
          nf, counts, offs, cap, orows, &ctl->nover, &ctl->need);
      ILLUM_LAUNCH_CHECK();
      ILLUM_HIP(hipMemcpyAsync(&nnz, sp->d_colptr + nf, 8, hipMemcpyDeviceToHost, s));
      ILLUM_HIP(hipMemcpyAsync(&h, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
      sync(ctx);
    };
    scan_and_read();
    if (qs.paged && h.nover_pg != 0) {
      // Sources that ran out of pool pages wrote no lists (zero counts):
      // they run again with the rerun shape (their rays are already
      // counted), and the counts are scanned again.  A paged rerun that runs
      // out of pages again lists the source in the pool-overflow rows
      // (overwritten below) and fails.
      const int nov = h.nover_pg;
      if (nov < 0 || nov > nf) throw Error(ILLUM_ERR_HIP, "form factors: bad page-overflow count");
      paged_overflow(nov, nf);
      FFArgs af = base_args(nov, 1, qf);
      af.rows = over_rows;
      af.over_rows = orows;
      af.over_n = &ctl->nover_pg2;
      af.row_counts = rcounts;
      af.pool = pool;
      af.pool_n = &ctl->pool_n;
      af.pool_cap = cap;
      af.vcounts = vcounts;
      af.voffs = voffs;
      af.vpool = vpool;
      af.vpool_n = &ctl->vpool_n;
      af.vpool_cap = vcap;
      {
        StageTimer st(ctx, "ff_visibility_rerun");
        launch_f<<<nov, FF_BLOCK, lds_f, s>>>(af);
        ILLUM_LAUNCH_CHECK();
      }
      // the pool-overflow count and need accumulate: cleared for the rescan
      ILLUM_HIP(hipMemsetAsync(&ctl->nover, 0, sizeof(int), s));
      ILLUM_HIP(hipMemsetAsync(&ctl->need, 0, sizeof(unsigned long long), s));
      scan_and_read();
    }
    rerun_overflow(h.nover_pg2);
    if (h.status) throw Error(ILLUM_ERR_NAN, "computed a NaN form factor");
    if (nnz < 0 || (unsigned long long)nnz != h.pool_n)
      throw Error(ILLUM_ERR_HIP, "form factors: inconsistent pool count");
    if (keep_vis && (int64_t)h.vpool_n > vcap) {
      // the raw visibility sets did not fit: the whole pass again
      if (pass > 0) throw Error(ILLUM_ERR_HIP, "form factors: pool overflow after resizing");
      vcap = std::max<int64_t>(vcap, (int64_t)h.vpool_n + (int64_t)h.vpool_n / 8 + 4096);
      cap = std::max<int64_t>(cap, (int64_t)h.pool_n + (int64_t)h.pool_n / 8 + 4096);
      continue;
    }
    if (h.nover > 0) {
      // Second pool: the overflowing sources again, their segments at
      // offsets cap + (position in pool2); counts, CSR row lengths, NaN
      // status and rays were all recorded by the first pass.
      if (h.nover > nf || (int64_t)h.need > nnz)
        throw Error(ILLUM_ERR_HIP, "form factors: inconsistent overflow count");
      cap2 = (int64_t)h.need;
      pool2 = (int32_t *)ctx_workspace(ctx, WS_FF_POOL2, 4 * (size_t)std::max<int64_t>(1, cap2));
      FFArgs a2 = base_args(h.nover, 1, qf);
      a2.rows = orows;
      a2.over_n = &ctl->nover_pg2;  // over_rows: free again
      a2.pool = pool2;
      a2.pool_n = &ctl->pool2_n;
      a2.pool_cap = cap2;
      a2.off_base = cap;
      {
        StageTimer st(ctx, "ff_visibility_rerun");
        launch_f<<<h.nover, FF_BLOCK, lds_f, s>>>(a2);
        ILLUM_LAUNCH_CHECK();
      }
      ILLUM_HIP(hipMemcpyAsync(&h, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
      sync(ctx);
      rerun_overflow(h.nover_pg2);
      if ((int64_t)h.pool2_n != cap2)
        throw Error(ILLUM_ERR_HIP, "form factors: second pool count differs");
    }
    break;
  }
  bvh->ff_nnz_last = nnz;
  if (nnz >= ((int64_t)1 << 40)) throw Error(ILLUM_ERR_CAPACITY, "nnz out of range");
  sp->nnz = nnz;
  sp->has_csc = true;
  sp->d_rowidx = palloc<int32_t>(ctx, nnz);
  sp->d_cval = palloc<double>(ctx, nnz);
  const unsigned g4 = (unsigned)((nf + 3) / 4);
  {
    StageTimer st(ctx, "ff_csc");
    ff_fill_list_kernel<<<g4, 256, 0, s>>>(nf, pool, cap, pool2, offs, sp->d_colptr,
                                           (const float4 *)mesh->d_pr4, (const float4 *)mesh->d_n4,
                                           mesh->d_area, d_alb, albedo, sp->d_rowidx, sp->d_cval);
    ILLUM_LAUNCH_CHECK();
  }
  if (keep_vis) {
    // the raw visibility sets, laid out by their own offsets
    sp->d_vptr = palloc<int64_t>(ctx, nf + 1);
    scan_counts(ctx, vcounts, nf, sp->d_vptr);
    sp->vnnz = (int64_t)h.vpool_n;
    sp->d_vidx = palloc<int32_t>(ctx, std::max<int64_t>(1, sp->vnnz));
    list_copy_kernel<<<g4, 256, 0, s>>>(nf, vpool, voffs, sp->d_vptr, sp->d_vidx);
    ILLUM_LAUNCH_CHECK();
  }
  sp->d_rowptr = palloc<int64_t>(ctx, nf + 1);
  sp->d_colidx = palloc<int32_t>(ctx, nnz + CSR_PAD);
  sp->d_rval = palloc<double>(ctx, nnz + CSR_PAD);
  ILLUM_HIP(hipMemsetAsync(sp->d_colidx + nnz, 0, 4 * CSR_PAD, s));
  ILLUM_HIP(hipMemsetAsync(sp->d_rval + nnz, 0, 8 * CSR_PAD, s));
  {
    // CSR by a sparse transpose of the CSC entries: scatter to rows through
    // atomic cursors (the unsorted rows in the transpose workspace), then
    // order short rows by rank in LDS and long rows through a bitmap.
    StageTimer st(ctx, "ff_csr");
    scan_counts(ctx, rcounts, nf, sp->d_rowptr);
    const size_t tjb = ((size_t)4 * nnz + 255) & ~(size_t)255;
    char *tmp = (char *)ctx_workspace(ctx, WS_FF_MASK_T, tjb + 8 * (size_t)std::max<int64_t>(1, nnz));
    int32_t *tj = (int32_t *)tmp;
    double *tv = (double *)(tmp + tjb);
    ILLUM_HIP(hipMemsetAsync(counts, 0, 4 * nf, s));  // the CSC counts are spent: cursors
    csr_scatter_kernel<<<g4, 256, 0, s>>>(nf, sp->d_colptr, sp->d_rowidx, sp->d_cval,
                                          sp->d_rowptr, counts, tj, tv);
    ILLUM_LAUNCH_CHECK();
    // ILLUM_FF_CSR_SORT_MAX (0..1024) lowers the rank sort's row limit, so
    // more rows take the bitmap path (for comparing the two).
    int64_t sort_max = CSR_SORT_MAX;
    if (const char *e = getenv("ILLUM_FF_CSR_SORT_MAX"))
      sort_max = std::max<int64_t>(0, std::min<int64_t>(CSR_SORT_MAX, atol(e)));
    csr_sort_kernel<<<g4, 256, 0, s>>>(nf, sort_max, sp->d_rowptr, tj, tv, sp->d_colidx,
                                       sp->d_rval);
    ILLUM_LAUNCH_CHECK();
    if (h.maxrow > sort_max) {
      // bitmap windows of up to 15,872 words, 62 KB beside the kernel's
      // static LDS within the default 64 KB (507,904 sources: one window for
      // any row of a body that size or smaller); ILLUM_FF_CSR_WINDOW (words)
      // narrows them, so that rows take several windows (results are
      // identical)
      int wwords = std::max(1, std::min(body_dims(mesh).nwmax, 15872));
      if (const char *e = getenv("ILLUM_FF_CSR_WINDOW"))
        wwords = std::max(1, std::min(wwords, atoi(e)));
      csr_long_kernel<<<(unsigned)nf, 256, 4 * (size_t)wwords, s>>>(
          sort_max, wwords, sp->d_rowptr, tj, (const float4 *)mesh->d_pr4,
          (const float4 *)mesh->d_n4, mesh->d_area, d_alb, albedo, sp->d_colidx, sp->d_rval);
      ILLUM_LAUNCH_CHECK();
    }
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
  const int layout = ff_resolve(bvh, layout_choice);
  const int stack_slots = ff_stack_slots(bvh, layout);
  const QShape qs = ff_qshape(body_dims(mesh), stack_slots);
  auto launch = ff_kernel<false>(layout, qs.paged);
  const size_t lds = ff_prepare(mesh, (const void *)launch, stack_slots, qs);
  // kept lists of the sampled rows, 64 entries per row (a row that does not
  // fit is counted and not written)
  const int64_t cap = 64 * (int64_t)nrows;
  DBuf<int32_t> counts(nrows), pool(cap);
  DBuf<int64_t> offs(nrows);
  DBuf<unsigned long long> pool_n(1);
  DBuf<int> status(1), over_n(1);
  DBuf<int32_t> over_rows(nrows);
  hipEvent_t e0, e1;
  ILLUM_HIP(hipEventCreate(&e0));
  ILLUM_HIP(hipEventCreate(&e1));
  FFArgs a = ff_args(bvh, nrows, jstride, stack_slots, qs, 1e-5, nullptr, 1.0);
  a.status = status.p;
  a.clus = cluster_cull() ? (const float4 *)mesh->d_clus : nullptr;
  a.counts = counts.p;
  a.offs = offs.p;
  a.pool = pool.p;
  a.pool_n = pool_n.p;
  a.pool_cap = cap;
  a.over_rows = over_rows.p;  // sources out of pool pages are timed, not rerun
  a.over_n = over_n.p;
  float best = 1e30f;
  for (int rep = 0; rep < 2; ++rep) {  // first launch also loads the code object
    ILLUM_HIP(hipMemsetAsync(pool_n.p, 0, 8, s));
    ILLUM_HIP(hipMemsetAsync(over_n.p, 0, 4, s));
    ILLUM_HIP(hipEventRecord(e0, s));
    launch<<<nrows, FF_BLOCK, lds, s>>>(a);
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
// instantiation: out[0..n) of {rays, pair tests, box tests, triangle tests
// (per active lane), wavefront walks, wavefront node visits, wavefront
// triangle tests, walks with a fixed permutation class} (the last four: the
// wavefront walks only).  The assembly kernel itself carries no counters;
// this is a separate launch, and writes no lists.
void form_factor_work(illum_bvh *bvh, double offset, int64_t *out, int n) {
  check_bvh_bodies(bvh);
  illum_mesh *mesh = bvh->mesh;
  illum_ctx *ctx = mesh->ctx;
  hipStream_t s = stream_of(ctx);
  const int64_t nf = mesh->g.nf;
  const int layout = ff_layout(bvh);
  const int stack_slots = ff_stack_slots(bvh, layout);
  const QShape qs = ff_qshape(body_dims(mesh), stack_slots);
  auto launch = ff_kernel<true>(layout, qs.paged);
  const size_t lds = ff_prepare(mesh, (const void *)launch, stack_slots, qs);
  DBuf<int> status(1), over_n(1);
  DBuf<int32_t> over_rows(nf);
  ILLUM_HIP(hipMemsetAsync(over_n.p, 0, 4, stream_of(ctx)));
  DBuf<unsigned long long> stats(8);
  ILLUM_HIP(hipMemsetAsync(stats.p, 0, 64, s));
  FFArgs a = ff_args(bvh, (int)nf, 1, stack_slots, qs, offset, nullptr, 1.0);
  a.status = status.p;
  a.stats = stats.p;
  a.clus = cluster_cull() ? (const float4 *)mesh->d_clus : nullptr;
  a.over_rows = over_rows.p;  // sources out of pool pages are counted, not rerun
  a.over_n = over_n.p;
  launch<<<(int)nf, FF_BLOCK, lds, s>>>(a);
  ILLUM_LAUNCH_CHECK();
  unsigned long long hs[8];
  ILLUM_HIP(hipMemcpyAsync(hs, stats.p, 64, hipMemcpyDeviceToHost, s));
  sync(ctx);
  for (int k = 0; k < n && k < 8; ++k) out[k] = (int64_t)hs[k];
}

}  // namespace illum
