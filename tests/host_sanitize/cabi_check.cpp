// The C ABI's host code under AddressSanitizer and UndefinedBehaviorSanitizer
// (the library built by `make -C illum_amd/csrc asan`: host code sanitized,
// device code as usual).  Runs on a GPU: every family of entry points on a
// small mesh (form factors, horizons, direct insolation, both radiosity
// methods at S = 1 and S = 128 with FP32 remainders, the no-allocation solve,
// conduction steps, runs and held-flux runs, the time loop, k-NN transfer,
// the renderer, the conductionQ drop-in), then the error paths (null
// pointers, bad sizes, unknown methods, a zero pivot), each of which must
// return its status code.  Any out-of-bounds host access, use after free or
// undefined operation ends the run with a report; a wrong status prints the
// call and makes the exit status nonzero.
#include "../../include/illum_hip.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

static int fails = 0;
#define OK(call)                                                                       \
  do {                                                                                 \
    const int st_ = (call);                                                            \
    if (st_ != ILLUM_OK) {                                                             \
      std::fprintf(stderr, "cabi_check: %s -> %d (%s)\n", #call, st_, illum_last_error()); \
      ++fails;                                                                         \
    }                                                                                  \
  } while (0)
#define STATUS(call, want)                                                             \
  do {                                                                                 \
    const int st_ = (call);                                                            \
    if (st_ != (want)) {                                                               \
      std::fprintf(stderr, "cabi_check: %s -> %d, expected %d\n", #call, st_, (int)(want)); \
      ++fails;                                                                         \
    }                                                                                  \
  } while (0)

int main(int argc, char **argv) {
  illum_ctx *ctx = nullptr;
  OK(illum_ctx_create(0, &ctx));
  if (!ctx) return 2;
  OK(illum_ctx_set_profiling(ctx, 1));

  // ---- mesh, BVH, form factors ---------------------------------------------
  int64_t nv = 0, nf = 0;
  OK(illum_synth_sizes(4, &nv, &nf));
  std::vector<double> V(3 * nv);
  std::vector<int32_t> F(3 * nf);
  OK(illum_synth_ellipsoid(4, 13, 11, 9, 40, 0.05, 0.3, 0.2, 1.5, 3, V.data(), F.data()));
  illum_mesh *mesh = nullptr;
  OK(illum_mesh_create(ctx, V.data(), nv, F.data(), nf, nullptr, 1.0, &mesh));
  illum_bvh *bvh = nullptr;
  OK(illum_bvh_build(mesh, 0, &bvh));  // autotuned leaf size and walk
  int nt = 0;
  std::vector<int> ls(64), lay(64);
  std::vector<double> tms(64);
  OK(illum_bvh_tuning(bvh, 64, ls.data(), lay.data(), tms.data(), &nt));
  int64_t tsz[4] = {0, 0, 0, 0};
  OK(illum_bvh_export(bvh, tsz, nullptr, nullptr, nullptr, nullptr));
  if (tsz[2] != nf || tsz[3] != 1) {
    std::fprintf(stderr, "cabi_check: exported tree sizes %lld triangles, %lld bodies\n",
                 (long long)tsz[2], (long long)tsz[3]);
    ++fails;
  }
  std::vector<float> tnodes(8 * tsz[0]), twide(16 * tsz[1]), ttris(12 * tsz[2]);
  std::vector<int32_t> troots(3 * tsz[3]);
  OK(illum_bvh_export(bvh, tsz, tnodes.data(), twide.data(), ttris.data(), troots.data()));
  STATUS(illum_bvh_export(bvh, nullptr, nullptr, nullptr, nullptr, nullptr), ILLUM_ERR_ARG);
  std::vector<double> alb(nf, 0.12);
  illum_sparse *Fm = nullptr;
  OK(illum_form_factors_ex(bvh, 0.0, alb.data(), 1e-5, ILLUM_FF_KEEP_VISIBILITY, &Fm));
  int64_t n = 0, nnz = 0, vnnz = 0, rays = 0;
  OK(illum_sparse_info(Fm, &n, &nnz));
  OK(illum_sparse_visibility_nnz(Fm, &vnnz));
  OK(illum_sparse_rays(Fm, &rays));
  std::vector<int64_t> cp(n + 1), rp(n + 1), vp(n + 1);
  std::vector<int32_t> ri(nnz), ci(nnz), vi(vnnz);
  std::vector<double> cv(nnz), rv(nnz);
  OK(illum_sparse_copy_csc(Fm, cp.data(), ri.data(), cv.data()));
  OK(illum_sparse_copy_csr(Fm, rp.data(), ci.data(), rv.data()));
  OK(illum_sparse_copy_visibility_lists(Fm, vp.data(), vi.data()));
  std::vector<int64_t> work(8);
  OK(illum_form_factors_work_ex(bvh, 1e-5, work.data(), 8));
  illum_sparse *Fr = nullptr;  // the same matrix through the CSR entry point
  OK(illum_sparse_create_csr(ctx, n, nnz, rp.data(), ci.data(), rv.data(), &Fr));

  // ---- horizons, direct insolation ----------------------------------------
  const int nphi = 33;
  std::vector<double> hz(nphi * nf);
  OK(illum_horizons(bvh, nullptr, nphi, 1e-3, 1e-5, hz.data(), 0));
  std::vector<double> disk(200);
  OK(illum_fib_spiral(disk.data(), 100));
  const int S = 128;
  std::vector<double> suns(3 * S);
  for (int s = 0; s < S; ++s) {
    const double a = 2 * M_PI * s / S;
    suns[3 * s] = 2.2739e8 * std::cos(a);
    suns[3 * s + 1] = 2.2739e8 * std::sin(a);
    suns[3 * s + 2] = 4.0e7;
  }
  std::vector<double> E(nf * S);
  OK(illum_direct_radiosity(mesh, hz.data(), nphi, suns.data(), S, disk.data(), 100, E.data(), 0));
  for (double &e : E) e += 1.0;  // every column nonzero

  // ---- radiosity: both methods, S = 1 and S = 128 --------------------------
  std::vector<double> B(nf * S), res(201);
  int it = 0;
  for (int method : {ILLUM_RAD_JACOBI_FP64, ILLUM_RAD_MIXED}) {
    OK(illum_radiosity_solve(Fm, E.data(), 1, method, 1e-10, 200, B.data(), res.data(), &it, 0));
    OK(illum_radiosity_solve(Fm, E.data(), S, method, 1e-8, 200, B.data(), res.data(), &it, 0));
    OK(illum_radiosity_solve(Fr, E.data(), S, method, 1e-11, 200, B.data(), res.data(), &it, 0));
  }
  STATUS(illum_radiosity_solve(Fm, E.data(), S, ILLUM_RAD_MIXED, 1e-14, 1, B.data(), res.data(),
                               &it, 0),
         ILLUM_ERR_NOT_CONVERGED);
  {  // no-allocation solve on device arrays
    void *dE = nullptr, *dB = nullptr;
    OK(illum_dev_alloc(ctx, 8 * nf * S, &dE));
    OK(illum_dev_alloc(ctx, 8 * nf * S, &dB));
    OK(illum_dev_copy(ctx, dE, E.data(), 8 * nf * S, 0));  // host to device
    OK(illum_radiosity_solve_noalloc(Fm, (const double *)dE, S, ILLUM_RAD_MIXED, 1e-8, 100,
                                     (double *)dB, res.data(), &it,
                                     ILLUM_IN_ON_DEVICE | ILLUM_OUT_ON_DEVICE));
    OK(illum_dev_copy(ctx, B.data(), dB, 8 * nf * S, 1));  // device to host
    OK(illum_dev_free(ctx, dE));
    OK(illum_dev_free(ctx, dB));
  }

  // ---- conduction ------------------------------------------------------------
  const int ng = 2;
  const int64_t goff[3] = {0, 40, 40 + 57};
  std::vector<double> z(97), ti(97, 200.0), rhoc(97, 1.19177e6);
  for (int k = 0; k < 40; ++k) z[k] = 0.5 * (k + 1) * (k + 1) / 1600.0;
  for (int k = 0; k < 57; ++k) z[40 + k] = 0.5 * (k + 1) * (k + 1) / 3249.0;
  std::vector<int32_t> gof(nf);
  std::vector<double> tis(nf);
  for (int64_t f = 0; f < nf; ++f) { gof[f] = (int32_t)(f % ng); tis[f] = 0.5 + (f % 7) * 0.1; }
  illum_thermal *th = nullptr;
  for (int prec : {ILLUM_THERM_FP64, ILLUM_THERM_MIXED}) {
    OK(illum_thermal_create(ctx, nf, ng, goff, z.data(), ti.data(), rhoc.data(), gof.data(),
                            tis.data(), 0.95, 230.0, nullptr, prec, &th));
    std::vector<double> Q(4 * nf, 300.0);
    OK(illum_thermal_step(th, 60.0, Q.data(), 0));
    OK(illum_thermal_run(th, 60.0, Q.data(), 4, 0));
    OK(illum_thermal_run(th, 60.0, Q.data(), 6, ILLUM_Q_HELD));
    std::vector<double> T(58 * nf), Ts(nf), Fs(nf), rad(nf, 1.0);
    OK(illum_thermal_get(th, T.data(), Ts.data(), Fs.data(), 0));
    OK(illum_thermal_get_radiosity(th, rad.data(), ILLUM_ACCUMULATE));
    OK(illum_thermal_destroy(th));
  }

  // ---- coupled time loop ---------------------------------------------------
  OK(illum_thermal_create(ctx, nf, 1, goff, z.data(), ti.data(), rhoc.data(), nullptr, nullptr,
                          0.95, 230.0, nullptr, ILLUM_THERM_MIXED, &th));
  {
    void *dh = nullptr, *dd = nullptr;
    OK(illum_dev_alloc(ctx, 8 * nphi * nf, &dh));
    OK(illum_dev_alloc(ctx, 8 * 200, &dd));
    OK(illum_dev_copy(ctx, dh, hz.data(), 8 * nphi * nf, 0));
    OK(illum_dev_copy(ctx, dd, disk.data(), 8 * 200, 0));
    for (int eager : {0, ILLUM_TL_EAGER}) {
      illum_timeloop *tl = nullptr;
      OK(illum_timeloop_create(Fm, mesh, th, (const double *)dh, nphi, (const double *)dd, 100,
                               ILLUM_RAD_MIXED, 1e-8, 0, eager, &tl));
      OK(illum_timeloop_run(tl, suns.data(), 12, 60.0, 0));
      int64_t steps = 0;
      int prods = 0, maxp = 0, unc = 0;
      double mr = 0;
      OK(illum_timeloop_stats(tl, &steps, &prods, &maxp, &unc, &mr));
      std::vector<double> tE(nf), tB(nf);
      OK(illum_timeloop_get(tl, tE.data(), tB.data(), 0));
      OK(illum_timeloop_destroy(tl));
    }
    OK(illum_dev_free(ctx, dh));
    OK(illum_dev_free(ctx, dd));
  }
  OK(illum_thermal_destroy(th));

  // ---- k-NN transfer, renderer, closest hits, conductionQ -----------------
  {
    std::vector<double> src(3 * 500), tgt(3 * 300), Xs(2 * 500), Xt(2 * 300), as(500, 1.0),
        at(300, 2.0), dist(300 * 4);
    for (int i = 0; i < 1500; ++i) src[i] = std::sin(0.37 * i) * 5;
    for (int i = 0; i < 900; ++i) tgt[i] = std::cos(0.53 * i) * 5;
    for (int i = 0; i < 1000; ++i) Xs[i] = i * 0.25;
    std::vector<int32_t> idx(300 * 4);
    OK(illum_knn(ctx, src.data(), 500, tgt.data(), 300, 4, idx.data(), dist.data(), 0));
    for (int mode = 0; mode < 3; ++mode)
      OK(illum_mesh_transfer(ctx, src.data(), 500, tgt.data(), 300, Xs.data(), 2, mode, 4,
                             as.data(), at.data(), Xt.data()));
    std::vector<int32_t> fac(64 * 48);
    std::vector<double> val(64 * 48), data(nf, 1.0);
    OK(illum_render(bvh, 0, 64, 48, 30.0, 20.0, data.data(), 0.0, 2.0, fac.data(), val.data(), 0));
    std::vector<float> o(3 * 100, 0.0f), d(3 * 100);
    for (int i = 0; i < 300; ++i) d[i] = (float)std::sin(1.3 * i);
    std::vector<int32_t> hit(100);
    std::vector<float> t(100);
    OK(illum_bvh_closest_hit(bvh, o.data(), d.data(), 100, hit.data(), t.data()));
    std::vector<double> Tz(40, 230.0);
    double Tsurf = 230.0, Fsurf = 0.0;
    OK(illum_conductionQ(40, z.data(), 60.0, 100.0, 120.0, Tz.data(), ti.data(), rhoc.data(),
                         0.95, &Tsurf, 0.0, &Fsurf));
  }

  // ---- error paths ---------------------------------------------------------
  illum_mesh *none = nullptr;
  STATUS(illum_mesh_create(ctx, nullptr, nv, F.data(), nf, nullptr, 1.0, &none), ILLUM_ERR_ARG);
  STATUS(illum_radiosity_solve(Fm, E.data(), 1, 77, 1e-8, 10, B.data(), nullptr, nullptr, 0),
         ILLUM_ERR_ARG);
  STATUS(illum_radiosity_solve(Fm, nullptr, 1, ILLUM_RAD_MIXED, 1e-8, 10, B.data(), nullptr,
                               nullptr, 0),
         ILLUM_ERR_ARG);
  STATUS(illum_horizons(bvh, nullptr, 1, 1e-3, 1e-5, hz.data(), 0), ILLUM_ERR_ARG);
  STATUS(illum_knn(ctx, nullptr, 500, nullptr, 300, 4, nullptr, nullptr, 0), ILLUM_ERR_ARG);
  {
    // a two-layer column whose first pivot b(1) is exactly 0 (tridag.for:10-12):
    // k = 1, alpha_1 = gamma_1 = 1/2, Tr = 1 and an emissivity for which
    // (2 eps sigma) Tr^3 rounds to -0.6, so bn = 4 and b(1) = 2 - 4/2 = 0
    const double sig = 5.6704e-8;
    double e = -0.3 / sig;
    for (int k = 0; k < 200; ++k) {
      const double brad = ((2.0 * e) * sig) * ((1.0 * 1.0) * 1.0);
      if (brad == -0.6) break;
      e = std::nextafter(e, brad > -0.6 ? -INFINITY : INFINITY);
    }
    std::vector<double> Tz = {1.0, 1.0}, zz = {0.5, 1.0}, tt = {1.0, 1.0}, rr = {1.0, 1.0};
    double Tsurf = 1.0, Fsurf = 0.0;
    STATUS(illum_conductionQ(2, zz.data(), 1.0, 10.0, 20.0, Tz.data(), tt.data(), rr.data(), e,
                             &Tsurf, 0.0, &Fsurf),
           ILLUM_ERR_ZERO_PIVOT);
  }
  STATUS(illum_mesh_from_obj(ctx, "/nonexistent/none.obj", 1.0, &none), ILLUM_ERR_IO);

  // ---- profile, tuning, teardown ------------------------------------------
  std::vector<double> pms(64);
  std::vector<int64_t> pl(64);
  std::vector<char> names(64 * 64);
  int len = 0, np = 0;
  OK(illum_ctx_get_profile(ctx, 64, pms.data(), pl.data(), names.data(), (int)names.size(), &np));
  std::vector<char> tun(1 << 16);
  OK(illum_ctx_tuning(ctx, tun.data(), (int)tun.size(), &len));
  OK(illum_sparse_destroy(Fr));
  OK(illum_sparse_destroy(Fm));
  OK(illum_bvh_destroy(bvh));
  OK(illum_mesh_destroy(mesh));
  OK(illum_ctx_destroy(ctx));
  std::printf("cabi_check: %lld facets, nnz %lld, %d failures\n", (long long)nf, (long long)nnz,
              fails);
  (void)argc;
  (void)argv;
  return fails ? 1 : 0;
}
