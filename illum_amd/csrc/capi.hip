// C ABI of libillum_hip: object lifetime, host<->device staging, error
// translation.  Every entry point catches illum::Error / std::exception and
// returns a status code with the message kept in a thread-local string.
#include "hip_util.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>

namespace illum {

// Implemented in the kernel translation units.
double time_visibility_sample(illum_bvh *bvh, int max_sources, int layout);
illum_sparse *form_factors(illum_bvh *bvh, double albedo, const double *albedo_vec,
                           double offset, int flags);
void render(illum_bvh *bvh, int mode, int width, int height, double az, double el,
            const double *data, double lo, double hi, int32_t *facet, double *value,
            int flags);
void form_factor_work(illum_bvh *bvh, double offset, int64_t *out, int n);
void bvh_closest_hit(illum_bvh *bvh, const float *orig, const float *dir, int64_t nray,
                     int32_t *hit, float *t);
void horizons(illum_bvh *bvh, illum_bvh *occl, int nphi, double theta_eps, double offset,
              double *out, int flags);
void direct_radiosity(illum_mesh *mesh, const double *hz, int nphi, const double *sun, int S,
                      const double *disk, int ndisk, double *E, int flags);
void radiosity_solve(illum_sparse *F, const double *E, int S, int method, double rel_tol,
                     int max_iter, double *B, double *resid, int *iters, int flags,
                     void *fixed_scratch, size_t fixed_cap, SolveStats *queue_stats);
void *radiosity_noalloc_scratch(int device);
illum_thermal *thermal_create(illum_ctx *ctx, int64_t nf, int ngrid, const int64_t *grid_off,
                              const double *z, const double *ti, const double *rhoc,
                              const int32_t *grid_of_facet, const double *ti_scale,
                              double emissivity, double T0, const double *T0_full,
                              int precision);
void thermal_run(illum_thermal *th, double dt, const double *Q, int nsteps, int flags,
                 double Fgeo, double *Tsurf_series);
void thermal_get(illum_thermal *th, double *T, double *Tsurf, double *Fsurf, int flags);
void thermal_get_radiosity(illum_thermal *th, double *out, int flags);
void thermal_free_impl(illum_thermal *th);

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

template <class Fn>
static int guarded(Fn &&fn) {
  return guarded_call(std::forward<Fn>(fn));
}

void collect_profile(illum_ctx *ctx) {
  for (auto &p : ctx->pending) {
    float ms = 0;
    hipEvent_t e0 = (hipEvent_t)p.second.first, e1 = (hipEvent_t)p.second.second;
    if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
      auto &st = ctx->profile[p.first];
      st.ms += ms;
      st.launches += 1;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  ctx->pending.clear();
}

void sync(illum_ctx *ctx) {
  ILLUM_HIP(hipStreamSynchronize(stream_of(ctx)));
  if (!ctx->pending.empty()) collect_profile(ctx);
}

// Exclusive scan of int32 counts into int64 offsets, reduce-then-scan over
// chunks of SCAN_CHUNK counts: block totals, a one-block scan of the totals
// (which also writes out[n]), then each block rescans its chunk from its
// offset.  Every thread handles SCAN_PER consecutive counts.
constexpr int SCAN_PER = 16;
constexpr int SCAN_CHUNK = 256 * SCAN_PER;

// Block-wide exclusive scan of one int64 per thread (256 threads); *total
// receives the block sum.
__device__ int64_t block_exscan256_i64(int64_t v, int64_t *s_wave, int64_t *total) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int64_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[wid] = x;
  __syncthreads();
  int64_t off = 0, tot = 0;
  for (int w = 0; w < 4; ++w) {
    const int64_t sw = s_wave[w];
    if (w < wid) off += sw;
    tot += sw;
  }
  __syncthreads();
  *total = tot;
  return off + x - v;
}

__global__ __launch_bounds__(256) void scan_reduce_kernel(const int32_t *__restrict__ counts,
                                                          int64_t n, int64_t *__restrict__ part) {
  __shared__ int64_t s_wave[4];
  const int64_t a = (int64_t)blockIdx.x * SCAN_CHUNK + (int64_t)threadIdx.x * SCAN_PER;
  int64_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k)
    if (a + k < n) sum += counts[a + k];
  int64_t total;
  block_exscan256_i64(sum, s_wave, &total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// One block: part[0..nb) becomes its exclusive scan; out_n = the grand total.
__global__ __launch_bounds__(256) void scan_part_kernel(int64_t *__restrict__ part, int nb,
                                                        int64_t *__restrict__ out_n) {
  __shared__ int64_t s_wave[4];
  const int per = (nb + 255) / 256;
  const int a = threadIdx.x * per, b = min(nb, a + per);
  int64_t sum = 0;
  for (int i = a; i < b; ++i) sum += part[i];
  int64_t total;
  int64_t run = block_exscan256_i64(sum, s_wave, &total);
  for (int i = a; i < b; ++i) {
    const int64_t v = part[i];
    part[i] = run;
    run += v;
  }
  if (threadIdx.x == 0) *out_n = total;
}

__global__ __launch_bounds__(256) void scan_apply_kernel(const int32_t *__restrict__ counts,
                                                         int64_t n,
                                                         const int64_t *__restrict__ part,
                                                         int64_t *__restrict__ out) {
  __shared__ int64_t s_wave[4];
  const int64_t a = (int64_t)blockIdx.x * SCAN_CHUNK + (int64_t)threadIdx.x * SCAN_PER;
  int32_t c[SCAN_PER];
  int64_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    c[k] = a + k < n ? counts[a + k] : 0;
    sum += c[k];
  }
  int64_t total;
  int64_t run = part[blockIdx.x] + block_exscan256_i64(sum, s_wave, &total);
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    if (a + k < n) out[a + k] = run;
    run += c[k];
  }
}

void scan_counts(illum_ctx *ctx, const int32_t *counts, int64_t n, int64_t *offsets) {
  hipStream_t s = stream_of(ctx);
  const int nb = (int)std::max<int64_t>(1, (n + SCAN_CHUNK - 1) / SCAN_CHUNK);
  int64_t *part = (int64_t *)ctx_workspace(ctx, WS_SCAN_PART, 8 * (size_t)nb);
  scan_reduce_kernel<<<nb, 256, 0, s>>>(counts, n, part);
  ILLUM_LAUNCH_CHECK();
  scan_part_kernel<<<1, 256, 0, s>>>(part, nb, offsets + n);
  ILLUM_LAUNCH_CHECK();
  scan_apply_kernel<<<nb, 256, 0, s>>>(counts, n, part, offsets);
  ILLUM_LAUNCH_CHECK();
}

static void set_device(illum_ctx *ctx) { ILLUM_HIP(hipSetDevice(ctx->device)); }

static void destroy_mesh(illum_mesh *m) {
  dfree(m->d_pr4); dfree(m->d_n4); dfree(m->d_area); dfree(m->d_frame); dfree(m->d_clus);
  dfree(m->d_boff); dfree(m->d_body); dfree(m->d_dq);
  delete m;
}
static void destroy_bvh(illum_bvh *b) {
  dfree(b->d_nodes); dfree(b->d_tris); dfree(b->d_wide); dfree(b->d_wide4); dfree(b->d_roots);
  delete b;
}
void destroy_sparse(illum_sparse *s) {
  // the context's reuse pool takes the blocks (no hipFree, no device sync)
  for (void *p : {(void *)s->d_colptr, (void *)s->d_rowidx, (void *)s->d_cval,
                  (void *)s->d_rowptr, (void *)s->d_colidx, (void *)s->d_rval,
                  (void *)s->d_vptr, (void *)s->d_vidx, (void *)s->d_nrays})
    pool_free(s->ctx, p);
  delete s;
}
void destroy_thermal(illum_thermal *t) {
  dfree(t->d_nz); dfree(t->d_grid); dfree(t->d_s2); dfree(t->d_goff); dfree(t->d_T);
  dfree(t->d_Qprev); dfree(t->d_F); dfree(t->d_work); dfree(t->d_work32); dfree(t->d_Qdev);
  dfree(t->d_status); dfree(t->d_fidx); dfree(t->d_pos); dfree(t->d_wgrid);
  thermal_free_impl(t);
  delete t;
}

// Device copy of the body partition (HostGeometry::body_off).
static void upload_bodies(illum_mesh *m) {
  hipStream_t s = stream_of(m->ctx);
  const int64_t nf = m->g.nf;
  if (m->g.body_off.size() < 2) m->g.body_off = {0, nf};
  const int nb = (int)m->g.body_off.size() - 1;
  std::vector<int32_t> off(m->g.body_off.begin(), m->g.body_off.end());
  dfree(m->d_boff);
  dfree(m->d_body);
  m->d_boff = nullptr;
  m->d_body = nullptr;
  m->nbody = nb;
  m->d_boff = dalloc<int32_t>(nb + 1);
  ILLUM_HIP(hipMemcpyAsync(m->d_boff, off.data(), 4 * (nb + 1), hipMemcpyHostToDevice, s));
  if (nb > 1) {
    std::vector<int32_t> body(nf);
    for (int b = 0; b < nb; ++b)
      for (int64_t f = off[b]; f < off[b + 1]; ++f) body[f] = b;
    m->d_body = dalloc<int32_t>(nf);
    ILLUM_HIP(hipMemcpyAsync(m->d_body, body.data(), 4 * nf, hipMemcpyHostToDevice, s));
  }
  sync(m->ctx);
}

static illum_mesh *upload_mesh(illum_ctx *ctx, HostGeometry &&g) {
  auto m = std::unique_ptr<illum_mesh, void (*)(illum_mesh *)>(new illum_mesh(), destroy_mesh);
  m->ctx = ctx;
  m->g = std::move(g);
  const int64_t nf = m->g.nf;
  std::vector<float> pr4(4 * nf), n4(4 * nf);
  for (int64_t f = 0; f < nf; ++f) {
    for (int c = 0; c < 3; ++c) {
      pr4[4 * f + c] = m->g.cen[3 * f + c];
      n4[4 * f + c] = m->g.nrm[3 * f + c];
    }
    pr4[4 * f + 3] = m->g.rad[f];
    n4[4 * f + 3] = 0.0f;
  }
  hipStream_t s = stream_of(ctx);
  m->d_pr4 = dalloc<float>(4 * nf);
  m->d_n4 = dalloc<float>(4 * nf);
  m->d_area = dalloc<double>(nf);
  m->d_frame = dalloc<double>(9 * nf);
  ILLUM_HIP(hipMemcpyAsync(m->d_pr4, pr4.data(), 16 * nf, hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(m->d_n4, n4.data(), 16 * nf, hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(m->d_area, m->g.area.data(), 8 * nf, hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(m->d_frame, m->g.frame.data(), 72 * nf, hipMemcpyHostToDevice, s));
  const std::vector<float> clus = cull_clusters(m->g);
  m->nclus = (int64_t)(clus.size() / 4);
  m->d_clus = dalloc<float>(clus.size());
  ILLUM_HIP(hipMemcpyAsync(m->d_clus, clus.data(), 4 * clus.size(), hipMemcpyHostToDevice, s));
  upload_bodies(m.get());
  // the direct-insolation queue: 8 open pairs per facet, 4,096 to 2^20
  m->dq_cap = std::min<int64_t>(std::max<int64_t>(8 * nf, 4096), (int64_t)1 << 20);
  m->d_dq = dalloc<int32_t>(m->dq_cap + 2);
  ILLUM_HIP(hipMemsetAsync(m->d_dq, 0, 8, s));
  sync(ctx);
  return m.release();
}

static illum_ctx *default_ctx() {
  static std::mutex mu;
  static illum_ctx *ctx = nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (!ctx) {
    auto c = std::unique_ptr<illum_ctx>(new illum_ctx());
    c->device = 0;
    ILLUM_HIP(hipSetDevice(0));
    c->noalloc_scratch = radiosity_noalloc_scratch(0);
    hipStream_t s;
    ILLUM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    c->stream = s;
    ctx = c.release();
  }
  return ctx;
}

std::vector<hipStream_t> fork_streams(illum_ctx *ctx, int n) {
  if (!ctx->fork_event) {
    hipEvent_t e;
    ILLUM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ctx->fork_event = e;
  }
  while ((int)ctx->aux_streams.size() < n) {
    hipStream_t s;
    ILLUM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    ctx->aux_streams.push_back(s);
    hipEvent_t e;
    ILLUM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ctx->aux_events.push_back(e);
  }
  ILLUM_HIP(hipEventRecord((hipEvent_t)ctx->fork_event, stream_of(ctx)));
  std::vector<hipStream_t> out(n);
  for (int k = 0; k < n; ++k) {
    out[k] = (hipStream_t)ctx->aux_streams[k];
    ILLUM_HIP(hipStreamWaitEvent(out[k], (hipEvent_t)ctx->fork_event, 0));
  }
  return out;
}

void join_streams(illum_ctx *ctx, int n) {
  for (int k = 0; k < n; ++k) {
    ILLUM_HIP(hipEventRecord((hipEvent_t)ctx->aux_events[k], (hipStream_t)ctx->aux_streams[k]));
    ILLUM_HIP(hipStreamWaitEvent(stream_of(ctx), (hipEvent_t)ctx->aux_events[k], 0));
  }
}

constexpr size_t POOL_MAX_BYTES = (size_t)8 << 30;

void *pool_alloc(illum_ctx *ctx, size_t bytes) {
  if (bytes == 0) bytes = 1;
  auto it = ctx->pool.lower_bound(bytes);
  if (it != ctx->pool.end() && it->first <= bytes + bytes / 2) {
    void *p = it->second;
    ctx->pool_bytes -= it->first;
    ctx->pool.erase(it);
    return p;
  }
  void *p = nullptr;
  ILLUM_HIP(hipMalloc(&p, bytes));
  ctx->pool_cap[p] = bytes;
  return p;
}

void pool_free(illum_ctx *ctx, void *p) {
  if (!p) return;
  auto c = ctx->pool_cap.find(p);
  if (c == ctx->pool_cap.end()) {
    (void)hipFree(p);
    return;
  }
  if (ctx->pool_bytes + c->second > POOL_MAX_BYTES) {
    (void)hipFree(p);
    ctx->pool_cap.erase(c);
    return;
  }
  ctx->pool.emplace(c->second, p);
  ctx->pool_bytes += c->second;
}

void *ctx_workspace(illum_ctx *ctx, int slot, size_t bytes) {
  if ((int)ctx->workspace.size() <= slot) ctx->workspace.resize(slot + 1, {nullptr, 0});
  auto &w = ctx->workspace[slot];
  if (w.second < bytes) {
    ILLUM_HIP(hipStreamSynchronize(stream_of(ctx)));
    if (w.first) ILLUM_HIP(hipFree(w.first));
    w.first = nullptr;
    w.second = 0;
    ILLUM_HIP(hipMalloc(&w.first, bytes));
    w.second = bytes;
  }
  return w.first;
}

}  // namespace illum

using namespace illum;

#define CHECK_ARG(cond, msg) \
  do { if (!(cond)) throw Error(ILLUM_ERR_ARG, msg); } while (0)

extern "C" {

const char *illum_last_error(void) { return g_last_error.c_str(); }
int illum_version(void) { return 1; }

int illum_ctx_create(int device, illum_ctx **out) {
  return guarded([&] {
    CHECK_ARG(out, "null output");
    int n = 0;
    ILLUM_HIP(hipGetDeviceCount(&n));
    CHECK_ARG(device >= 0 && device < n, "device index out of range");
    auto c = std::unique_ptr<illum_ctx>(new illum_ctx());
    c->device = device;
    ILLUM_HIP(hipSetDevice(device));
    // the no-allocation solve's scratch address on this device (resolved
    // once per device, with it current)
    c->noalloc_scratch = radiosity_noalloc_scratch(device);
    hipStream_t s;
    ILLUM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    c->stream = s;
    *out = c.release();
  });
}

int illum_ctx_destroy(illum_ctx *ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipStreamSynchronize(stream_of(ctx));
    collect_profile(ctx);
    for (void *a : ctx->aux_streams) (void)hipStreamSynchronize((hipStream_t)a);
    (void)hipStreamDestroy(stream_of(ctx));
    for (void *a : ctx->aux_streams) (void)hipStreamDestroy((hipStream_t)a);
    for (void *e : ctx->aux_events) (void)hipEventDestroy((hipEvent_t)e);
    if (ctx->fork_event) (void)hipEventDestroy((hipEvent_t)ctx->fork_event);
    for (auto &w : ctx->workspace) (void)hipFree(w.first);
    for (auto &b : ctx->pool) (void)hipFree(b.second);
    delete ctx;
  });
}

int illum_ctx_synchronize(illum_ctx *ctx) {
  return guarded([&] { CHECK_ARG(ctx, "null context"); sync(ctx); });
}

int illum_ctx_set_profiling(illum_ctx *ctx, int enable) {
  return guarded([&] { CHECK_ARG(ctx, "null context"); ctx->profiling = enable != 0; });
}

int illum_ctx_reset_profile(illum_ctx *ctx) {
  return guarded([&] {
    CHECK_ARG(ctx, "null context");
    sync(ctx);
    ctx->profile.clear();
  });
}

int illum_ctx_get_profile(illum_ctx *ctx, int cap, double *total_ms, int64_t *launches,
                          char *names, int names_cap, int *n) {
  return guarded([&] {
    CHECK_ARG(ctx && n, "null argument");
    sync(ctx);
    int k = 0, pos = 0;
    for (auto &p : ctx->profile) {
      if (k < cap) {
        if (total_ms) total_ms[k] = p.second.ms;
        if (launches) launches[k] = p.second.launches;
        if (names) {
          const int len = (int)p.first.size() + 1;
          if (pos + len <= names_cap) {
            std::memcpy(names + pos, p.first.c_str(), len);
            pos += len;
          }
        }
      }
      ++k;
    }
    *n = k;
  });
}

int illum_ctx_tuning(illum_ctx *ctx, char *out, int cap, int *len) {
  return guarded([&] {
    CHECK_ARG(ctx && len, "null argument");
    std::string r;
    for (auto &t : ctx->tuned)
      r += t.key + " " + std::to_string(t.choice) + " " + std::to_string(t.best_ms) + " " +
           t.trials + "\n";
    *len = (int)r.size();
    if (out && cap > 0) {
      const int n = std::min<int>(cap - 1, (int)r.size());
      std::memcpy(out, r.data(), n);
      out[n] = 0;
    }
  });
}

int illum_dev_alloc(illum_ctx *ctx, int64_t bytes, void **out) {
  return guarded([&] {
    CHECK_ARG(ctx && out && bytes >= 0, "bad argument");
    set_device(ctx);
    *out = dalloc<char>((size_t)bytes);
  });
}

int illum_dev_free(illum_ctx *ctx, void *p) {
  return guarded([&] {
    CHECK_ARG(ctx, "null context");
    set_device(ctx);
    sync(ctx);
    if (p) ILLUM_HIP(hipFree(p));
  });
}

int illum_dev_copy(illum_ctx *ctx, void *dst, const void *src, int64_t bytes, int kind) {
  return guarded([&] {
    CHECK_ARG(ctx && dst && src && bytes >= 0 && kind >= 0 && kind <= 2, "bad argument");
    set_device(ctx);
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                     : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    ILLUM_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, k, stream_of(ctx)));
    sync(ctx);
  });
}

int illum_synth_sizes(int subdiv, int64_t *nv, int64_t *nf) {
  return guarded([&] { synth_sizes(subdiv, nv, nf); });
}

int illum_synth_ellipsoid(int subdiv, double a, double b, double c, int ncraters, double dmin,
                          double dmax, double wmin, double wmax, uint64_t seed, double *V,
                          int32_t *F) {
  return guarded([&] {
    CHECK_ARG(V && F, "null output");
    synth_ellipsoid(subdiv, a, b, c, ncraters, dmin, dmax, wmin, wmax, seed, V, F);
  });
}

int illum_mesh_create(illum_ctx *ctx, const double *V, int64_t nv, const int32_t *F,
                      int64_t nf, const double *VN, double vertex_divisor, illum_mesh **out) {
  return guarded([&] {
    CHECK_ARG(ctx && out, "null argument");
    set_device(ctx);
    *out = upload_mesh(ctx, make_geometry(V, nv, F, nf, VN, vertex_divisor));
  });
}

int illum_mesh_from_obj(illum_ctx *ctx, const char *path, double vertex_divisor,
                        illum_mesh **out) {
  return guarded([&] {
    CHECK_ARG(ctx && path && out, "null argument");
    set_device(ctx);
    std::vector<double> V, VN;
    std::vector<int32_t> F;
    read_obj(path, V, F, VN);
    const int64_t nv = (int64_t)(V.size() / 3);
    // vertex normals are indexed by vertex index (obj_util.cpp:71-85): used
    // when there is one for every vertex the faces reference
    int64_t vmax = 0;
    for (int32_t v : F) vmax = std::max<int64_t>(vmax, v);
    const double *vn = (int64_t)(VN.size() / 3) > vmax ? VN.data() : nullptr;
    *out = upload_mesh(ctx, make_geometry(V.data(), nv, F.data(), (int64_t)(F.size() / 3), vn,
                                          vertex_divisor));
  });
}

int illum_mesh_destroy(illum_mesh *mesh) {
  return guarded([&] { if (mesh) destroy_mesh(mesh); });
}

int illum_mesh_num_faces(const illum_mesh *mesh, int64_t *nf) {
  return guarded([&] { CHECK_ARG(mesh && nf, "null argument"); *nf = mesh->g.nf; });
}

int illum_mesh_set_bodies(illum_mesh *mesh, const int64_t *body_face_off, int nbody) {
  return guarded([&] {
    CHECK_ARG(mesh && body_face_off && nbody >= 1, "bad argument");
    const int64_t nf = mesh->g.nf;
    CHECK_ARG(body_face_off[0] == 0 && body_face_off[nbody] == nf,
              "body offsets must start at 0 and end at the facet count");
    for (int b = 0; b < nbody; ++b)
      CHECK_ARG(body_face_off[b + 1] > body_face_off[b], "every body needs at least one facet");
    set_device(mesh->ctx);
    mesh->g.body_off.assign(body_face_off, body_face_off + nbody + 1);
    upload_bodies(mesh);
  });
}

int illum_mesh_geometry(const illum_mesh *mesh, float *centroid, float *normal, float *radius,
                        double *area, double *frame) {
  return guarded([&] {
    CHECK_ARG(mesh, "null mesh");
    const auto &g = mesh->g;
    if (centroid) std::memcpy(centroid, g.cen.data(), 4 * g.cen.size());
    if (normal) std::memcpy(normal, g.nrm.data(), 4 * g.nrm.size());
    if (radius) std::memcpy(radius, g.rad.data(), 4 * g.rad.size());
    if (area) std::memcpy(area, g.area.data(), 8 * g.area.size());
    if (frame) std::memcpy(frame, g.frame.data(), 8 * g.frame.size());
  });
}

int illum_mesh_triangles(const illum_mesh *mesh, float *tri9) {
  return guarded([&] {
    CHECK_ARG(mesh && tri9, "null argument");
    std::memcpy(tri9, mesh->g.tri.data(), 4 * mesh->g.tri.size());
  });
}

static illum_bvh *make_bvh(illum_mesh *mesh, int leaf_size) {
  illum_ctx *ctx = mesh->ctx;
  auto b = std::unique_ptr<illum_bvh, void (*)(illum_bvh *)>(new illum_bvh(), destroy_bvh);
  b->mesh = mesh;
  b->h = build_bvh(mesh->g, leaf_size);
  hipStream_t s = stream_of(ctx);
  b->d_nodes = dalloc<float>(b->h.nodes.size());
  b->d_tris = dalloc<float>(b->h.tris.size());
  b->d_wide = dalloc<float>(b->h.wide.size());
  ILLUM_HIP(hipMemcpyAsync(b->d_wide, b->h.wide.data(), 4 * b->h.wide.size(),
                           hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(b->d_nodes, b->h.nodes.data(), 4 * b->h.nodes.size(),
                           hipMemcpyHostToDevice, s));
  ILLUM_HIP(hipMemcpyAsync(b->d_tris, b->h.tris.data(), 4 * b->h.tris.size(),
                           hipMemcpyHostToDevice, s));
  b->d_wide4 = dalloc<float>(b->h.wide4.size());
  ILLUM_HIP(hipMemcpyAsync(b->d_wide4, b->h.wide4.data(), 4 * b->h.wide4.size(),
                           hipMemcpyHostToDevice, s));
  b->nbody = (int)b->h.root_thr.size();
  b->body_off = mesh->g.body_off;
  std::vector<int32_t> roots(4 * b->nbody, 0);
  for (int k = 0; k < b->nbody; ++k) {
    roots[4 * k] = b->h.root_thr[k];
    roots[4 * k + 1] = b->h.root_wide[k];
    roots[4 * k + 2] = b->h.root_wide4[k];
  }
  b->d_roots = dalloc<int32_t>(roots.size());
  ILLUM_HIP(hipMemcpyAsync(b->d_roots, roots.data(), 4 * roots.size(), hipMemcpyHostToDevice, s));
  sync(ctx);
  return b.release();
}

int illum_bvh_build(illum_mesh *mesh, int leaf_size, illum_bvh **out) {
  return guarded([&] {
    CHECK_ARG(mesh && out, "null argument");
    CHECK_ARG(leaf_size >= 0, "leaf_size must be >= 0 (0 = autotune)");
    set_device(mesh->ctx);
    if (leaf_size > 0) {
      *out = make_bvh(mesh, leaf_size);
      return;
    }
    // Autotune: time the visibility kernel on a strided sample of up to 8192
    // sources for each candidate leaf size and traversal layout (two-wide,
    // four-wide, wavefront packets over the two-wide and four-wide trees) and
    // keep the fastest tree and layout.
    static const int cand[] = {2, 3, 4, 6, 8};
    std::unique_ptr<illum_bvh, void (*)(illum_bvh *)> best(nullptr, destroy_bvh);
    double best_ms = 0;
    int best_layout = 2;
    int leaves[32], layouts[32];
    double times[32];
    int n = 0;
    for (int c : cand) {
      std::unique_ptr<illum_bvh, void (*)(illum_bvh *)> b(make_bvh(mesh, c), destroy_bvh);
      bool improved = false;
      for (int layout : {2, 4, 3, 5}) {
        const double ms = time_visibility_sample(b.get(), 8192, layout);
        leaves[n] = c;
        layouts[n] = layout;
        times[n] = ms;
        ++n;
        if (n == 1 || ms < best_ms) {
          best_ms = ms;
          best_layout = layout;
          improved = true;
        }
      }
      if (improved) best = std::move(b);
    }
    for (int k = 0; k < n; ++k) {
      best->tune_leaf[k] = leaves[k];
      best->tune_layout[k] = layouts[k];
      best->tune_ms[k] = times[k];
    }
    best->ntuned = n;
    best->ff_layout = best_layout;
    *out = best.release();
  });
}

int illum_bvh_tuning(const illum_bvh *bvh, int cap, int *leaf_sizes, int *layouts, double *ms,
                     int *n) {
  return guarded([&] {
    CHECK_ARG(bvh && n, "null argument");
    for (int k = 0; k < bvh->ntuned && k < cap; ++k) {
      if (leaf_sizes) leaf_sizes[k] = bvh->tune_leaf[k];
      if (layouts) layouts[k] = bvh->tune_layout[k];
      if (ms) ms[k] = bvh->tune_ms[k];
    }
    *n = bvh->ntuned;
  });
}

int illum_bvh_destroy(illum_bvh *bvh) {
  return guarded([&] { if (bvh) destroy_bvh(bvh); });
}

int illum_bvh_info(const illum_bvh *bvh, int64_t *nodes, int *depth, int *leaf_size) {
  return guarded([&] {
    CHECK_ARG(bvh, "null bvh");
    if (nodes) *nodes = bvh->h.nnodes;
    if (depth) *depth = bvh->h.depth;
    if (leaf_size) *leaf_size = bvh->h.leaf_size;
  });
}

int illum_bvh_export(const illum_bvh *bvh, int64_t *sizes, float *nodes, float *wide,
                     float *tris, int32_t *roots) {
  return guarded([&] {
    CHECK_ARG(bvh && sizes, "null argument");
    const illum::HostBVH &h = bvh->h;
    const int64_t nb = (int64_t)h.root_thr.size();
    sizes[0] = h.nnodes;
    sizes[1] = h.ninner;
    sizes[2] = (int64_t)h.tris.size() / 12;
    sizes[3] = nb;
    if (nodes) std::memcpy(nodes, h.nodes.data(), 4 * h.nodes.size());
    if (wide) std::memcpy(wide, h.wide.data(), 4 * h.wide.size());
    if (tris) std::memcpy(tris, h.tris.data(), 4 * h.tris.size());
    if (roots)
      for (int64_t b = 0; b < nb; ++b) {
        roots[3 * b] = h.root_thr[b];
        roots[3 * b + 1] = b < (int64_t)h.root_wide.size() ? h.root_wide[b] : -1;
        roots[3 * b + 2] = b < (int64_t)h.root_wide4.size() ? h.root_wide4[b] : -1;
      }
  });
}

int illum_bvh_closest_hit(illum_bvh *bvh, const float *orig, const float *dir, int64_t nray,
                          int32_t *hit, float *t) {
  return guarded([&] {
    CHECK_ARG(bvh && orig && dir && hit && t && nray >= 0, "bad argument");
    set_device(bvh->mesh->ctx);
    bvh_closest_hit(bvh, orig, dir, nray, hit, t);
  });
}

int illum_form_factors(illum_bvh *bvh, double albedo, const double *albedo_vec, double offset,
                       illum_sparse **out) {
  return illum_form_factors_ex(bvh, albedo, albedo_vec, offset, 0, out);
}

int illum_form_factors_ex(illum_bvh *bvh, double albedo, const double *albedo_vec,
                          double offset, int flags, illum_sparse **out) {
  return guarded([&] {
    CHECK_ARG(bvh && out, "null argument");
    CHECK_ARG(std::isfinite(offset), "offset must be finite");
    CHECK_ARG((flags & ~ILLUM_FF_KEEP_VISIBILITY) == 0, "unknown form-factor flags");
    set_device(bvh->mesh->ctx);
    *out = form_factors(bvh, albedo, albedo_vec, offset, flags);
  });
}

int illum_render(illum_bvh *bvh, int mode, int width, int height, double az_deg,
                 double el_deg, const double *data, double minval, double maxval,
                 int32_t *facet, double *value, int flags) {
  return guarded([&] {
    CHECK_ARG(bvh && (facet || value), "null argument");
    CHECK_ARG(!(minval > maxval), "minval must not exceed maxval");
    set_device(bvh->mesh->ctx);
    render(bvh, mode, width, height, az_deg, el_deg, data, minval, maxval, facet, value, flags);
  });
}

int illum_form_factors_work(illum_bvh *bvh, double offset, int64_t *out4) {
  return illum_form_factors_work_ex(bvh, offset, out4, 4);
}

int illum_form_factors_work_ex(illum_bvh *bvh, double offset, int64_t *out, int n) {
  return guarded([&] {
    CHECK_ARG(bvh && out && n >= 1 && n <= 8, "bad argument");
    CHECK_ARG(std::isfinite(offset), "offset must be finite");
    set_device(bvh->mesh->ctx);
    form_factor_work(bvh, offset, out, n);
  });
}

int illum_sparse_create_csr(illum_ctx *ctx, int64_t n, int64_t nnz, const int64_t *rowptr,
                            const int32_t *colidx, const double *val, illum_sparse **out) {
  return guarded([&] {
    CHECK_ARG(ctx && out && rowptr && n > 0 && nnz >= 0, "bad argument");
    CHECK_ARG(nnz == 0 || (colidx && val), "null index/value arrays");
    CHECK_ARG(rowptr[0] == 0 && rowptr[n] == nnz, "rowptr must start at 0 and end at nnz");
    for (int64_t i = 0; i < n; ++i)
      CHECK_ARG(rowptr[i + 1] >= rowptr[i], "rowptr must be non-decreasing");
    for (int64_t k = 0; k < nnz; ++k)
      CHECK_ARG(colidx[k] >= 0 && colidx[k] < n, "column index out of range");
    set_device(ctx);
    auto sp = std::unique_ptr<illum_sparse, void (*)(illum_sparse *)>(new illum_sparse(),
                                                                      destroy_sparse);
    sp->ctx = ctx;
    sp->n = n;
    sp->nnz = nnz;
    hipStream_t s = stream_of(ctx);
    sp->d_rowptr = palloc<int64_t>(ctx, n + 1);
    sp->d_colidx = palloc<int32_t>(ctx, nnz + CSR_PAD);
    sp->d_rval = palloc<double>(ctx, nnz + CSR_PAD);
    ILLUM_HIP(hipMemsetAsync(sp->d_colidx + nnz, 0, 4 * CSR_PAD, s));
    ILLUM_HIP(hipMemsetAsync(sp->d_rval + nnz, 0, 8 * CSR_PAD, s));
    ILLUM_HIP(hipMemcpyAsync(sp->d_rowptr, rowptr, 8 * (n + 1), hipMemcpyHostToDevice, s));
    if (nnz) {
      ILLUM_HIP(hipMemcpyAsync(sp->d_colidx, colidx, 4 * nnz, hipMemcpyHostToDevice, s));
      ILLUM_HIP(hipMemcpyAsync(sp->d_rval, val, 8 * nnz, hipMemcpyHostToDevice, s));
    }
    sync(ctx);
    *out = sp.release();
  });
}

int illum_sparse_destroy(illum_sparse *sp) {
  return guarded([&] { if (sp) destroy_sparse(sp); });
}

int illum_sparse_info(const illum_sparse *sp, int64_t *n, int64_t *nnz) {
  return guarded([&] {
    CHECK_ARG(sp, "null matrix");
    if (n) *n = sp->n;
    if (nnz) *nnz = sp->nnz;
  });
}

int illum_sparse_copy_csc(const illum_sparse *sp, int64_t *colptr, int32_t *rowidx,
                          double *val) {
  return guarded([&] {
    CHECK_ARG(sp, "null matrix");
    if (!sp->has_csc) throw Error(ILLUM_ERR_ARG, "matrix has no CSC form");
    hipStream_t s = stream_of(sp->ctx);
    if (colptr) ILLUM_HIP(hipMemcpyAsync(colptr, sp->d_colptr, 8 * (sp->n + 1), hipMemcpyDeviceToHost, s));
    if (rowidx && sp->nnz) ILLUM_HIP(hipMemcpyAsync(rowidx, sp->d_rowidx, 4 * sp->nnz, hipMemcpyDeviceToHost, s));
    if (val && sp->nnz) ILLUM_HIP(hipMemcpyAsync(val, sp->d_cval, 8 * sp->nnz, hipMemcpyDeviceToHost, s));
    sync(sp->ctx);
  });
}

int illum_sparse_copy_csr(const illum_sparse *sp, int64_t *rowptr, int32_t *colidx,
                          double *val) {
  return guarded([&] {
    CHECK_ARG(sp, "null matrix");
    hipStream_t s = stream_of(sp->ctx);
    if (rowptr) ILLUM_HIP(hipMemcpyAsync(rowptr, sp->d_rowptr, 8 * (sp->n + 1), hipMemcpyDeviceToHost, s));
    if (colidx && sp->nnz) ILLUM_HIP(hipMemcpyAsync(colidx, sp->d_colidx, 4 * sp->nnz, hipMemcpyDeviceToHost, s));
    if (val && sp->nnz) ILLUM_HIP(hipMemcpyAsync(val, sp->d_rval, 8 * sp->nnz, hipMemcpyDeviceToHost, s));
    sync(sp->ctx);
  });
}

int illum_sparse_visibility_words(const illum_sparse *sp, int64_t *words_per_row) {
  return guarded([&] {
    CHECK_ARG(sp && words_per_row, "null argument");
    if (!sp->d_vptr) throw Error(ILLUM_ERR_ARG, "matrix carries no visibility sets");
    *words_per_row = (sp->n + 31) / 32;
  });
}

int illum_sparse_visibility_nnz(const illum_sparse *sp, int64_t *nnz) {
  return guarded([&] {
    CHECK_ARG(sp && nnz, "null argument");
    if (!sp->d_vptr) throw Error(ILLUM_ERR_ARG, "matrix carries no visibility sets");
    *nnz = sp->vnnz;
  });
}

int illum_sparse_copy_visibility_lists(const illum_sparse *sp, int64_t *ptr, int32_t *idx) {
  return guarded([&] {
    CHECK_ARG(sp, "null matrix");
    if (!sp->d_vptr) throw Error(ILLUM_ERR_ARG, "matrix carries no visibility sets");
    hipStream_t s = stream_of(sp->ctx);
    if (ptr) ILLUM_HIP(hipMemcpyAsync(ptr, sp->d_vptr, 8 * (sp->n + 1), hipMemcpyDeviceToHost, s));
    if (idx && sp->vnnz)
      ILLUM_HIP(hipMemcpyAsync(idx, sp->d_vidx, 4 * sp->vnnz, hipMemcpyDeviceToHost, s));
    sync(sp->ctx);
  });
}

int illum_sparse_rays(const illum_sparse *sp, int64_t *rays) {
  return guarded([&] {
    CHECK_ARG(sp && rays, "null argument");
    if (!sp->d_nrays) throw Error(ILLUM_ERR_ARG, "matrix was not assembled by illum_form_factors");
    unsigned long long n = 0;
    ILLUM_HIP(hipMemcpy(&n, sp->d_nrays, 8, hipMemcpyDeviceToHost));
    *rays = (int64_t)n;
  });
}

int illum_sparse_copy_visibility(const illum_sparse *sp, uint32_t *bits) {
  return guarded([&] {
    CHECK_ARG(sp && bits, "null argument");
    if (!sp->d_vptr) throw Error(ILLUM_ERR_ARG, "matrix carries no visibility sets");
    // the lists expanded into bit rows on the host
    std::vector<int64_t> ptr(sp->n + 1);
    std::vector<int32_t> idx(std::max<int64_t>(1, sp->vnnz));
    hipStream_t s = stream_of(sp->ctx);
    ILLUM_HIP(hipMemcpyAsync(ptr.data(), sp->d_vptr, 8 * (sp->n + 1), hipMemcpyDeviceToHost, s));
    if (sp->vnnz)
      ILLUM_HIP(hipMemcpyAsync(idx.data(), sp->d_vidx, 4 * sp->vnnz, hipMemcpyDeviceToHost, s));
    sync(sp->ctx);
    const int64_t words = (sp->n + 31) / 32;
    std::memset(bits, 0, 4 * (size_t)(sp->n * words));
    for (int64_t j = 0; j < sp->n; ++j)
      for (int64_t k = ptr[j]; k < ptr[j + 1]; ++k)
        bits[j * words + (idx[k] >> 5)] |= 1u << (idx[k] & 31);
  });
}

int illum_horizons(illum_bvh *bvh, illum_bvh *occluder, int nphi, double theta_eps,
                   double offset, double *out, int flags) {
  return guarded([&] {
    CHECK_ARG(bvh, "null bvh");
    set_device(bvh->mesh->ctx);
    horizons(bvh, occluder, nphi, theta_eps, offset, out, flags);
  });
}

int illum_fib_spiral(double *xy, int n) {
  return guarded([&] {
    CHECK_ARG(xy && n > 0, "bad argument");
    // illum.cpp:468-483; golden ratio as arma::datum::gratio.
    const double TWO_PI = 2 * M_PI;
    const double gr = 1.6180339887498948482;
    double a = 0.0, da = (TWO_PI * (gr - 1)) / gr;
    double r = 0.0, dr = 1.0 / (n + 1);
    for (int i = 0; i < n; ++i) {
      xy[i] = r * std::cos(a);
      xy[n + i] = r * std::sin(a);
      a = std::fmod(a + da, TWO_PI);
      r += dr;
    }
  });
}

int illum_direct_radiosity(illum_mesh *mesh, const double *horizons_, int nphi,
                           const double *sun_km, int S, const double *disk_xy, int ndisk,
                           double *E, int flags) {
  return guarded([&] {
    CHECK_ARG(mesh, "null mesh");
    set_device(mesh->ctx);
    direct_radiosity(mesh, horizons_, nphi, sun_km, S, disk_xy, ndisk, E, flags);
  });
}

int illum_radiosity_solve(illum_sparse *F, const double *E, int S, int method, double rel_tol,
                          int max_iter, double *B, double *resid, int *iters, int flags) {
  return guarded([&] {
    CHECK_ARG(F, "null matrix");
    set_device(F->ctx);
    radiosity_solve(F, E, S, method, rel_tol, max_iter, B, resid, iters, flags, nullptr, 0,
                    nullptr);
  });
}

int illum_radiosity_solve_noalloc(illum_sparse *F, const double *E, int S, int method,
                                  double rel_tol, int max_iter, double *B, double *resid,
                                  int *iters, int flags) {
  return guarded([&] {
    CHECK_ARG(F, "null matrix");
    if (!F->ctx->noalloc_scratch) throw Error(ILLUM_ERR_HIP, "no-allocation scratch not resolved");
    set_device(F->ctx);
    radiosity_solve(F, E, S, method, rel_tol, max_iter, B, resid, iters, flags,
                    F->ctx->noalloc_scratch, RAD_NOALLOC_SCRATCH_BYTES, nullptr);
  });
}

int illum_thermal_create(illum_ctx *ctx, int64_t nf, int ngrid, const int64_t *grid_off,
                         const double *z, const double *ti, const double *rhoc,
                         const int32_t *grid_of_facet, const double *ti_scale,
                         double emissivity, double T0, const double *T0_full, int precision,
                         illum_thermal **out) {
  return guarded([&] {
    CHECK_ARG(ctx && out, "null argument");
    set_device(ctx);
    *out = thermal_create(ctx, nf, ngrid, grid_off, z, ti, rhoc, grid_of_facet, ti_scale,
                          emissivity, T0, T0_full, precision);
  });
}

int illum_thermal_destroy(illum_thermal *th) {
  return guarded([&] { if (th) destroy_thermal(th); });
}

int illum_thermal_info(const illum_thermal *th, int64_t *nf, int *nzmax) {
  return guarded([&] {
    CHECK_ARG(th, "null thermal model");
    if (nf) *nf = th->nf;
    if (nzmax) *nzmax = th->nzmax;
  });
}

int illum_thermal_step(illum_thermal *th, double dt, const double *Q, int flags) {
  return guarded([&] {
    CHECK_ARG(th, "null thermal model");
    set_device(th->ctx);
    thermal_run(th, dt, Q, 1, flags, 0.0, nullptr);
  });
}

int illum_thermal_run(illum_thermal *th, double dt, const double *Qseries, int nsteps,
                      int flags) {
  return guarded([&] {
    CHECK_ARG(th, "null thermal model");
    set_device(th->ctx);
    thermal_run(th, dt, Qseries, nsteps, flags, 0.0, nullptr);
  });
}

int illum_thermal_get(illum_thermal *th, double *T, double *Tsurf, double *Fsurf, int flags) {
  return guarded([&] {
    CHECK_ARG(th, "null thermal model");
    set_device(th->ctx);
    thermal_get(th, T, Tsurf, Fsurf, flags);
  });
}

int illum_thermal_get_radiosity(illum_thermal *th, double *out, int flags) {
  return guarded([&] {
    CHECK_ARG(th && out, "null argument");
    set_device(th->ctx);
    thermal_get_radiosity(th, out, flags);
  });
}

int illum_conductionQ(int nz, const double *z, double dt, double Qn, double Qnp1, double *T,
                      const double *ti, const double *rhoc, double emiss, double *Tsurf,
                      double Fgeotherm, double *Fsurf) {
  return guarded([&] {
    CHECK_ARG(nz >= 2 && z && T && ti && rhoc && Tsurf && Fsurf, "bad argument");
    illum_ctx *ctx = default_ctx();
    set_device(ctx);
    const int64_t off[2] = {0, nz};
    std::vector<double> T0(nz + 1);
    T0[0] = *Tsurf;
    for (int i = 0; i < nz; ++i) T0[i + 1] = T[i];
    illum_thermal *th = thermal_create(ctx, 1, 1, off, z, ti, rhoc, nullptr, nullptr, emiss,
                                       0.0, T0.data(), ILLUM_THERM_FP64);
    std::unique_ptr<illum_thermal, void (*)(illum_thermal *)> guard(th, destroy_thermal);
    ILLUM_HIP(hipMemcpy(th->d_Qprev, &Qn, 8, hipMemcpyHostToDevice));
    thermal_run(th, dt, &Qnp1, 1, 0, Fgeotherm, nullptr);
    std::vector<double> Tout(nz + 1);
    double F = 0;
    thermal_get(th, Tout.data(), nullptr, &F, 0);
    *Tsurf = Tout[0];
    for (int i = 0; i < nz; ++i) T[i] = Tout[i + 1];
    *Fsurf = F;
  });
}

}  // extern "C"
