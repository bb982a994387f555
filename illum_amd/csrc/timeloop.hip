// The coupled radiosity + thermal time loop of do_radiosity_task
// (illum_cli.cpp:403-456), device-resident and captured in HIP graphs.
//
// One step: E = direct insolation under the step's sun (illum.cpp:382-466)
// minus the thermal model's surface flux (the reference's rad = -F_surf +
// direct, illum_cli.cpp:411-423, summed inside the direct kernel's store),
// the radiosity solve (I - F) B = E, then one Crank-Nicolson step with Q = B
// (illum_cli.cpp:446).  The reference solves with a fixed number of
// Gauss-Seidel sweeps; here the solve stops on the device at a relative
// residual, within a fixed budget of products queued ahead (the ones after
// convergence return at once), so a step never waits on the host and the
// whole step is one captured graph.  The sun of step k is read through a
// device step counter, so one graph serves every step.  The loop owns E, B
// and the solve's scratch: a captured graph holds their addresses, which
// nothing else may move.
#include "hip_util.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace illum {

void direct_enqueue(illum_mesh *mesh, const double *hz, int nphi, const double *sun, int S,
                    const double *disk, int ndisk, double *E, const int *step,
                    const double *flux, hipStream_t s);
void radiosity_solve(illum_sparse *F, const double *E, int S, int method, double rel_tol,
                     int max_iter, double *B, double *resid, int *iters, int flags,
                     void *fixed_scratch, size_t fixed_cap, SolveStats *queue_stats);
size_t radiosity_scratch_bytes(int64_t nf, int S, int method, int max_iter);
void thermal_prepare(illum_thermal *th, double dt);
void thermal_enqueue(illum_thermal *th, double dt, const double *dQ, int nsteps);
void thermal_check(illum_thermal *th);

__global__ void set_int_kernel(int *p, int v) { *p = v; }
__global__ void step_advance_kernel(int *p) { *p = *p + 1; }
__global__ void stats_reset_kernel(SolveStats *st) {
  st->max_resid = 0.0;
  st->max_products = 0;
  st->unconverged = 0;
  st->solves = 0;
  st->pad = 0;
}

}  // namespace illum

using namespace illum;

struct illum_timeloop {
  illum_ctx *ctx = nullptr;
  illum_mesh *mesh = nullptr;
  illum_sparse *F = nullptr;
  illum_thermal *th = nullptr;
  const double *hz = nullptr;
  int nphi = 0;
  const double *disk = nullptr;
  int ndisk = 0;
  int method = ILLUM_RAD_MIXED;
  double rel_tol = 1e-8;
  int budget = 0;               // products queued per solve (0: not chosen yet)
  bool eager = false;           // queue kernels directly instead of replaying graphs
  double *d_E = nullptr, *d_B = nullptr;
  void *d_scr = nullptr;
  size_t scr_bytes = 0;
  double *d_sun = nullptr;
  int64_t sun_cap = 0;          // suns the buffer holds
  int *d_step = nullptr;
  SolveStats *d_stats = nullptr;
  int64_t steps_done = 0;
  // captured graphs of 1 and TL_GRAPH_STEPS steps, valid for graph_dt and
  // the current sun buffer
  hipGraphExec_t g1 = nullptr, gn = nullptr;
  double graph_dt = 0.0;
  const double *graph_sun = nullptr;
  int graph_budget = 0;
};

namespace {

constexpr int TL_GRAPH_STEPS = 8;

void drop_graphs(illum_timeloop *tl) {
  if (tl->g1) (void)hipGraphExecDestroy(tl->g1);
  if (tl->gn) (void)hipGraphExecDestroy(tl->gn);
  tl->g1 = tl->gn = nullptr;
}

void destroy_loop(illum_timeloop *tl) {
  if (tl->ctx) (void)hipStreamSynchronize(stream_of(tl->ctx));
  drop_graphs(tl);
  dfree(tl->d_E); dfree(tl->d_B); dfree(tl->d_sun); dfree(tl->d_step); dfree(tl->d_stats);
  if (tl->d_scr) (void)hipFree(tl->d_scr);
  delete tl;
}

// Queues one step on the context stream: no host work, no allocation.
void enqueue_step(illum_timeloop *tl, double dt) {
  hipStream_t s = stream_of(tl->ctx);
  direct_enqueue(tl->mesh, tl->hz, tl->nphi, tl->d_sun, 1, tl->disk, tl->ndisk, tl->d_E,
                 tl->d_step, tl->th->d_F, s);
  radiosity_solve(tl->F, tl->d_E, 1, tl->method, tl->rel_tol, tl->budget, tl->d_B, nullptr,
                  nullptr, ILLUM_IN_ON_DEVICE | ILLUM_OUT_ON_DEVICE, tl->d_scr, tl->scr_bytes,
                  tl->d_stats);
  thermal_enqueue(tl->th, dt, tl->d_B, 1);
  step_advance_kernel<<<1, 1, 0, s>>>(tl->d_step);
  ILLUM_LAUNCH_CHECK();
}

hipGraphExec_t capture(illum_timeloop *tl, int nsteps, double dt) {
  hipStream_t s = stream_of(tl->ctx);
  ILLUM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    for (int k = 0; k < nsteps; ++k) enqueue_step(tl, dt);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hipGraph_t g = nullptr;
  ILLUM_HIP(hipStreamEndCapture(s, &g));
  hipGraphExec_t e = nullptr;
  const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  ILLUM_HIP(err);
  return e;
}

// The solve's product budget when none was given: two more than a checked
// solve of the first step's right-hand side takes, at least 4 (a later
// solve that needs more is counted as unconverged and reported).
void choose_budget(illum_timeloop *tl) {
  hipStream_t s = stream_of(tl->ctx);
  set_int_kernel<<<1, 1, 0, s>>>(tl->d_step, 0);
  ILLUM_LAUNCH_CHECK();
  direct_enqueue(tl->mesh, tl->hz, tl->nphi, tl->d_sun, 1, tl->disk, tl->ndisk, tl->d_E,
                 tl->d_step, tl->th->d_F, s);
  int iters = 0;
  try {
    radiosity_solve(tl->F, tl->d_E, 1, tl->method, tl->rel_tol, 200, tl->d_B, nullptr, &iters,
                    ILLUM_IN_ON_DEVICE | ILLUM_OUT_ON_DEVICE, nullptr, 0, nullptr);
  } catch (const Error &e) {
    if (e.code != ILLUM_ERR_NOT_CONVERGED) throw;
    iters = 200;
  }
  tl->budget = std::max(4, iters + 2);
}

}  // namespace

#define CHECK_ARG(cond, msg) \
  do { if (!(cond)) throw Error(ILLUM_ERR_ARG, msg); } while (0)

extern "C" {

int illum_timeloop_create(illum_sparse *F, illum_mesh *mesh, illum_thermal *th,
                          const double *horizons, int nphi, const double *disk_xy, int ndisk,
                          int method, double rel_tol, int products, int flags,
                          illum_timeloop **out) {
  return guarded_call([&] {
    CHECK_ARG(F && mesh && th && horizons && disk_xy && out, "null argument");
    CHECK_ARG(F->ctx == mesh->ctx && th->ctx == mesh->ctx, "handles of different contexts");
    CHECK_ARG(F->n == mesh->g.nf && th->nf == mesh->g.nf, "matrix, mesh and thermal model sizes differ");
    CHECK_ARG(F->d_rowptr, "matrix has no CSR form");
    CHECK_ARG(nphi >= 2 && ndisk >= 1, "bad horizon or disk sizes");
    CHECK_ARG(method == ILLUM_RAD_JACOBI_FP64 || method == ILLUM_RAD_MIXED, "unknown radiosity method");
    CHECK_ARG(rel_tol >= 0 && products >= 0 && products <= 1000, "bad solve parameters");
    CHECK_ARG((flags & ~ILLUM_TL_EAGER) == 0, "unknown time-loop flags");
    illum_ctx *ctx = mesh->ctx;
    ILLUM_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<illum_timeloop, void (*)(illum_timeloop *)> tl(new illum_timeloop(),
                                                                   destroy_loop);
    tl->ctx = ctx;
    tl->mesh = mesh;
    tl->F = F;
    tl->th = th;
    tl->hz = horizons;
    tl->nphi = nphi;
    tl->disk = disk_xy;
    tl->ndisk = ndisk;
    tl->method = method;
    tl->rel_tol = rel_tol;
    tl->budget = products;
    // ILLUM_TL_EAGER=1 in the environment forces the eager form (the same
    // kernels, queued directly), for tools that do not follow graph launches
    const char *ev = getenv("ILLUM_TL_EAGER");
    tl->eager = (flags & ILLUM_TL_EAGER) || (ev && ev[0] == '1');
    const int64_t nf = mesh->g.nf;
    tl->d_E = dalloc<double>(nf);
    tl->d_B = dalloc<double>(nf);
    tl->d_step = dalloc<int>(1);
    tl->d_stats = dalloc<SolveStats>(1);
    *out = tl.release();
  });
}

int illum_timeloop_run(illum_timeloop *tl, const double *sun_km, int nsteps, double dt,
                       int flags) {
  return guarded_call([&] {
    CHECK_ARG(tl && sun_km && nsteps >= 1 && dt > 0, "bad argument");
    illum_ctx *ctx = tl->ctx;
    ILLUM_HIP(hipSetDevice(ctx->device));
    hipStream_t s = stream_of(ctx);
    // the suns, in the loop's own buffer (a graph holds its address)
    if (tl->sun_cap < nsteps) {
      ILLUM_HIP(hipStreamSynchronize(s));
      dfree(tl->d_sun);
      tl->d_sun = dalloc<double>(3 * (size_t)nsteps);
      tl->sun_cap = nsteps;
    }
    ILLUM_HIP(hipMemcpyAsync(tl->d_sun, sun_km, 24 * (size_t)nsteps,
                             (flags & ILLUM_IN_ON_DEVICE) ? hipMemcpyDeviceToDevice
                                                          : hipMemcpyHostToDevice, s));
    if (tl->budget == 0) choose_budget(tl);
    // owned solve scratch for the budget, and the thermal tables of dt (host
    // work: before any capture)
    const size_t need = radiosity_scratch_bytes(tl->mesh->g.nf, 1, tl->method, tl->budget);
    if (tl->scr_bytes < need) {
      ILLUM_HIP(hipStreamSynchronize(s));
      drop_graphs(tl);
      if (tl->d_scr) ILLUM_HIP(hipFree(tl->d_scr));
      tl->d_scr = nullptr;
      ILLUM_HIP(hipMalloc(&tl->d_scr, need));
      tl->scr_bytes = need;
    }
    thermal_prepare(tl->th, dt);
    set_int_kernel<<<1, 1, 0, s>>>(tl->d_step, 0);
    ILLUM_LAUNCH_CHECK();
    // one eager direct pass (its result is recomputed by step 0): tunes the
    // kernel's lanes per pair before any capture
    direct_enqueue(tl->mesh, tl->hz, tl->nphi, tl->d_sun, 1, tl->disk, tl->ndisk, tl->d_E,
                   tl->d_step, tl->th->d_F, s);
    ILLUM_LAUNCH_CHECK();
    stats_reset_kernel<<<1, 1, 0, s>>>(tl->d_stats);
    ILLUM_LAUNCH_CHECK();
    if (tl->eager) {
      for (int k = 0; k < nsteps; ++k) enqueue_step(tl, dt);
    } else {
      if (tl->graph_dt != dt || tl->graph_sun != tl->d_sun || tl->graph_budget != tl->budget)
        drop_graphs(tl);
      if (!tl->g1) {
        tl->g1 = capture(tl, 1, dt);
        tl->gn = capture(tl, TL_GRAPH_STEPS, dt);
        tl->graph_dt = dt;
        tl->graph_sun = tl->d_sun;
        tl->graph_budget = tl->budget;
      }
      int k = 0;
      for (; k + TL_GRAPH_STEPS <= nsteps; k += TL_GRAPH_STEPS) ILLUM_HIP(hipGraphLaunch(tl->gn, s));
      for (; k < nsteps; ++k) ILLUM_HIP(hipGraphLaunch(tl->g1, s));
    }
    tl->steps_done += nsteps;
    if ((flags & ILLUM_IN_ON_DEVICE) && (flags & ILLUM_ASYNC)) return;
    SolveStats h{};
    ILLUM_HIP(hipMemcpyAsync(&h, tl->d_stats, sizeof(h), hipMemcpyDeviceToHost, s));
    sync(ctx);
    thermal_check(tl->th);
    if (h.unconverged > 0)
      throw Error(ILLUM_ERR_NOT_CONVERGED,
                  std::to_string(h.unconverged) + " of " + std::to_string(h.solves) +
                      " solves ended above the tolerance within " + std::to_string(tl->budget) +
                      " products (largest residual " + std::to_string(h.max_resid) + ")");
  });
}

int illum_timeloop_stats(illum_timeloop *tl, int64_t *steps, int *products, int *max_products,
                         int *unconverged, double *max_resid) {
  return guarded_call([&] {
    CHECK_ARG(tl, "null time loop");
    ILLUM_HIP(hipSetDevice(tl->ctx->device));
    SolveStats h{};
    ILLUM_HIP(hipMemcpyAsync(&h, tl->d_stats, sizeof(h), hipMemcpyDeviceToHost,
                             stream_of(tl->ctx)));
    sync(tl->ctx);
    if (steps) *steps = tl->steps_done;
    if (products) *products = tl->budget;
    if (max_products) *max_products = h.max_products;
    if (unconverged) *unconverged = h.unconverged;
    if (max_resid) *max_resid = h.max_resid;
  });
}

int illum_timeloop_get(illum_timeloop *tl, double *E, double *B, int flags) {
  return guarded_call([&] {
    CHECK_ARG(tl, "null time loop");
    ILLUM_HIP(hipSetDevice(tl->ctx->device));
    hipStream_t s = stream_of(tl->ctx);
    const hipMemcpyKind k = (flags & ILLUM_OUT_ON_DEVICE) ? hipMemcpyDeviceToDevice
                                                         : hipMemcpyDeviceToHost;
    const size_t bytes = 8 * (size_t)tl->mesh->g.nf;
    if (E) ILLUM_HIP(hipMemcpyAsync(E, tl->d_E, bytes, k, s));
    if (B) ILLUM_HIP(hipMemcpyAsync(B, tl->d_B, bytes, k, s));
    sync(tl->ctx);
  });
}

int illum_timeloop_destroy(illum_timeloop *tl) {
  return guarded_call([&] { if (tl) destroy_loop(tl); });
}

}  // extern "C"
