// HIP helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdlib>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "internal.hpp"

#define ILLUM_HIP(call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::illum::Error(e_ == hipErrorOutOfMemory ? ILLUM_ERR_OOM : ILLUM_ERR_HIP,     \
                           std::string(#call) + ": " + hipGetErrorString(e_));           \
  } while (0)

#define ILLUM_LAUNCH_CHECK() ILLUM_HIP(hipGetLastError())

namespace illum {

// Runs fn, translating exceptions into a status code (the message kept for
// illum_last_error): the body of every C entry point.
template <class Fn>
int guarded_call(Fn &&fn) {
  try {
    fn();
    return ILLUM_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return ILLUM_ERR_OOM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return ILLUM_ERR_ARG;
  }
}

inline hipStream_t stream_of(const illum_ctx *ctx) { return (hipStream_t)ctx->stream; }

template <class T>
T *dalloc(size_t n) {
  if (n == 0) n = 1;
  void *p = nullptr;
  ILLUM_HIP(hipMalloc(&p, n * sizeof(T)));
  return (T *)p;
}

template <class T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

// RAII device buffer for temporaries inside one call.
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() {}
  explicit DBuf(size_t n_) : p(dalloc<T>(n_)), n(n_) {}
  ~DBuf() { dfree(p); }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  void reset(size_t n_) { dfree(p); p = dalloc<T>(n_); n = n_; }
};

// Records hipEvents around a named stage on the context stream when profiling
// is enabled; durations are collected at the next synchronisation point.
struct StageTimer {
  illum_ctx *ctx;
  std::string name;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  StageTimer(illum_ctx *c, const char *n, bool enable = true) : ctx(c), name(n) {
    if (ctx->profiling && enable) {
      ILLUM_HIP(hipEventCreate(&e0));
      ILLUM_HIP(hipEventCreate(&e1));
      ILLUM_HIP(hipEventRecord(e0, stream_of(ctx)));
    }
  }
  ~StageTimer() {
    if (ctx->profiling && e0) {
      if (hipEventRecord(e1, stream_of(ctx)) == hipSuccess)
        ctx->pending.push_back({name, {(void *)e0, (void *)e1}});
    }
  }
};

// Free every device buffer of a handle and the handle itself.
void destroy_sparse(illum_sparse *s);
void destroy_thermal(illum_thermal *t);

// Device scratch slot `slot` of the context, at least `bytes` long (grown by
// reallocation after the stream drains; contents are not preserved).  Slots
// live until the context is destroyed, so repeated calls (a time loop)
// allocate nothing; a context is used by one host thread at a time.
void *ctx_workspace(illum_ctx *ctx, int slot, size_t bytes);
enum WorkspaceSlot {
  WS_FF_POOL = 0,   // kept receivers of every source (form-factor assembly)
  WS_FF_MASK_T = 1, // unsorted CSR rows of the sparse transpose
  WS_FF_COUNTS = 2, // per-source counts and pool offsets, CSR row lengths
  WS_RAD_CTL = 3,   // solver control, residual history, column norms, partial sums
  WS_RAD_E = 4,     // right-hand sides, row-major
  WS_RAD_X0 = 5,    // iterates / residuals (nf x S each)
  WS_RAD_X1 = 6,
  WS_RAD_X2 = 7,
  WS_RAD_D32 = 8,   // FP16 corrections (ring)
  WS_SCAN_PART = 9, // scan_counts block totals
  WS_FF_VPOOL = 10, // raw visibility sets, when kept
  WS_FF_ALBEDO = 11,
  WS_FF_STATUS = 12,
  WS_FF_POOL2 = 13, // kept receivers of the sources that did not fit the pool
  WS_FF_ROWS = 14,  // those sources
  WS_FF_OVER = 15,  // sources of a paged launch that ran out of pool pages
};

// Size of the no-allocation solve's fixed per-GPU scratch array.
constexpr size_t RAD_NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;

// Running record of a time loop's queued radiosity solves (device).
struct SolveStats {
  double max_resid;  // largest final relative residual
  int max_products;  // most products a solve took
  int unconverged;   // solves that ended above the tolerance
  int solves;
  int pad;
};

// Device blocks from the context's reuse pool (see illum_ctx::pool): a free
// block of capacity in [bytes, 1.5 bytes] if there is one, else a new one.
// pool_free returns a block to the pool (or frees it past the pool's cap;
// blocks not from pool_alloc are freed).
void *pool_alloc(illum_ctx *ctx, size_t bytes);
void pool_free(illum_ctx *ctx, void *p);
template <class T>
T *palloc(illum_ctx *ctx, size_t n) {
  return (T *)pool_alloc(ctx, (n ? n : 1) * sizeof(T));
}

// Fork / join of independent launches: fork_streams makes n auxiliary
// streams wait for the work queued so far on the context stream and returns
// them; join_streams makes the context stream wait for everything queued on
// them since.  A call that forks joins before it returns.
std::vector<hipStream_t> fork_streams(illum_ctx *ctx, int n);
void join_streams(illum_ctx *ctx, int n);

// Autotuner: returns the cached choice for `key`, or times run(c) for each
// candidate c on the context stream (hipEvents, best of two launches) and
// keeps the fastest.  run(c) must give the same results for every c.  While
// the stream is being captured into a graph nothing is timed: `fallback` is
// returned and nothing cached.  ILLUM_NO_AUTOTUNE=1 returns `fallback`.
// may_time = false (callers that must not create events or allocate, such
// as the no-allocation solve) only looks the choice up.
template <class Run>
int autotune(illum_ctx *ctx, const std::string &key, const std::vector<int> &cands,
             int fallback, Run &&run, bool may_time = true) {
  auto it = ctx->tune_cache.find(key);
  if (it != ctx->tune_cache.end()) return it->second;
  const char *off = getenv("ILLUM_NO_AUTOTUNE");
  if ((off && off[0] == '1') || !may_time) return fallback;
  hipStream_t s = stream_of(ctx);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  ILLUM_HIP(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone) return fallback;
  hipEvent_t e0, e1;
  ILLUM_HIP(hipEventCreate(&e0));
  ILLUM_HIP(hipEventCreate(&e1));
  int best = fallback;
  float best_ms = 1e30f;
  std::string trials;
  for (int c : cands) {
    float ms = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {  // the first launch also loads the code
      ILLUM_HIP(hipEventRecord(e0, s));
      run(c);
      ILLUM_HIP(hipEventRecord(e1, s));
      ILLUM_HIP(hipEventSynchronize(e1));
      float t = 0;
      ILLUM_HIP(hipEventElapsedTime(&t, e0, e1));
      ms = std::min(ms, t);
    }
    trials += (trials.empty() ? "" : " ") + std::to_string(c) + ":" + std::to_string(ms);
    if (ms < best_ms) { best_ms = ms; best = c; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->tune_cache[key] = best;
  ctx->tuned.push_back({key, best, (double)best_ms, trials});
  return best;
}

// Drains pending stage events into the profile (after the stream is synced).
void collect_profile(illum_ctx *ctx);
void sync(illum_ctx *ctx);

// Exclusive scan of n int32 counts into n+1 int64 offsets (device), on stream.
void scan_counts(illum_ctx *ctx, const int32_t *counts, int64_t n, int64_t *offsets);

// Maps consecutive workgroup ids (dispatched round-robin over the 8 XCDs) to
// contiguous ranges per XCD, so each XCD's L2 serves spatially coherent
// work.  Bijective on [0, n).
__device__ __forceinline__ int xcd_remap(int b, int n) {
  const int x = b & 7, k = b >> 3;
  const int per = n >> 3, rem = n & 7;
  return x * per + (x < rem ? x : rem) + k;
}

// Maps workgroup ids so each XCD takes chunks of C consecutive ids in turn:
// XCD x (ids b = x mod 8, dispatched round-robin) gets chunks x, x + 8, ...
// of the id range.  Neighbouring work stays together within a chunk while
// every XCD draws from the whole range, so uneven work (a facet region with
// many rays) spreads over all XCDs.  Ids past the last whole round of 8
// chunks map to themselves.  Bijective on [0, n); C = 1 is the identity.
__device__ __forceinline__ int xcd_chunked(int b, int n, int C) {
  const int full = (n / (8 * C)) * 8 * C;
  if (b >= full) return b;
  const int x = b & 7, k = b >> 3;
  return ((k / C) * 8 + x) * C + k % C;
}

// Number of set bits of m in the lanes below this one (a lane's place among
// the ballot's set lanes): two mbcnt instructions, no mask constant.
__device__ __forceinline__ int rank_in_mask(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const unsigned lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

}  // namespace illum
