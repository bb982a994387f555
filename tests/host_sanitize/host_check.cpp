// Host-side checks under AddressSanitizer and UndefinedBehaviorSanitizer:
// the library's host code (mesh generator, OBJ reader, geometry, receiver
// clusters, BVH builds in every layout, single and multi-body) and the CPU
// oracle (every entry point, on a small mesh).  Built by the Makefile beside
// this file with -fsanitize=address,undefined -fno-sanitize-recover=all, so
// any out-of-bounds access, use after free, leak or undefined operation ends
// the run with a report and a nonzero status.  Consistency checks between
// the two (geometry bit for bit, BVH walks equal to brute force) make a wrong
// index fail too.  Usage: host_check <path to two_shapes.obj>
#include "internal.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

extern "C" {
int oracle_geometry(const double *V, int64_t nv, const int32_t *F, int64_t nf, const double *VN,
                    double divisor, float *tri9, float *cen, float *nrm, float *rad, double *area,
                    double *frame);
int oracle_closest_hit(const float *tri9, int64_t nf, const float *orig, const float *dir,
                       int64_t nray, int use_bvh, int nthreads, int32_t *hit, float *t);
int64_t oracle_visibility(const float *tri9, const float *cen, const float *nrm, const float *rad,
                          int64_t nf, double offset, int64_t j0, int64_t j1, int use_bvh,
                          int nthreads, uint32_t *bits);
int64_t oracle_form_factors(const float *cen, const float *nrm, const double *area, int64_t nf,
                            const uint32_t *bits, int64_t j0, int64_t j1, double albedo,
                            const double *albedo_vec, int64_t cap, int64_t *colptr,
                            int32_t *rowidx, double *val);
int64_t oracle_horizons(const float *cen, const float *nrm, const double *frame, int64_t j0,
                        int64_t j1, const float *occ9, int64_t nocc, int nphi, double theta_eps,
                        double offset, int use_bvh, int nthreads, double *out);
int oracle_fib_spiral(double *xy, int n);
int oracle_direct(const float *cen, const double *frame, int64_t nf, const double *hz, int nphi,
                  const double *sun, int S, const double *disk, int ndisk, int nthreads,
                  double *E);
int oracle_gauss_seidel(int64_t n, const int64_t *colptr, const int32_t *rowidx,
                        const double *val, const double *E, int max_sweeps, double rel_tol,
                        double *B, double *abs_resid);
int oracle_conductionQ(int nz, const double *z, double dt, double Qn, double Qnp1, double *T,
                       const double *ti, const double *rhoc, double emiss, double *Tsurf,
                       double Fgeotherm, double *Fsurf);
int oracle_thermal_run(int64_t nf, const int64_t *goff, const double *z, const double *ti,
                       const double *rhoc, const int32_t *grid, const double *ti_scale,
                       double emiss, int nzmax, double dt, int nsteps, const double *Q,
                       double *Qprev, double *T, double *Fsurf, double *Tsurf_series,
                       int nthreads);
int oracle_render(const float *tri9, const float *cen, int64_t nf, int mode, int width,
                  int height, double az, double el, int nthreads, int32_t *facet);
}

namespace illum {
void set_last_error(const std::string &) {}
}

static int fails = 0;
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "host_check: %s failed (line %d)\n", #c, __LINE__); \
      ++fails;                                                          \
    }                                                                   \
  } while (0)

template <class V, class T>
static bool same(const V &a, const T *b, size_t n) {
  return a.size() == n && std::memcmp(a.data(), b, n * sizeof(T)) == 0;
}

// The product's geometry against the oracle's, bit for bit.
static void check_geometry(const illum::HostGeometry &g, const double *V, int64_t nv,
                           const int32_t *F, const double *VN, double div) {
  const int64_t nf = g.nf;
  std::vector<float> tri(9 * nf), cen(3 * nf), nrm(3 * nf), rad(nf);
  std::vector<double> area(nf), frame(9 * nf);
  CHECK(oracle_geometry(V, nv, F, nf, VN, div, tri.data(), cen.data(), nrm.data(), rad.data(),
                        area.data(), frame.data()) == 0);
  CHECK(same(g.tri, tri.data(), tri.size()));
  CHECK(same(g.cen, cen.data(), cen.size()));
  CHECK(same(g.nrm, nrm.data(), nrm.size()));
  CHECK(same(g.rad, rad.data(), rad.size()));
  CHECK(same(g.area, area.data(), area.size()));
  CHECK(same(g.frame, frame.data(), frame.size()));
}

int main(int argc, char **argv) {
  using namespace illum;
  // ---- generator, geometry, clusters, BVH ----------------------------------
  int64_t nv = 0, nf = 0;
  synth_sizes(3, &nv, &nf);
  std::vector<double> V(3 * nv);
  std::vector<int32_t> F(3 * nf);
  synth_ellipsoid(3, 13.0, 11.0, 9.0, 30, 0.05, 0.3, 0.2, 1.5, 5, V.data(), F.data());
  HostGeometry g = make_geometry(V.data(), nv, F.data(), nf, nullptr, 1.0);
  CHECK(g.nf == nf);
  check_geometry(g, V.data(), nv, F.data(), nullptr, 1.0);
  const std::vector<float> cl = cull_clusters(g);
  CHECK(cl.size() == 4 * (size_t)((nf + 63) / 64));
  for (int leaf : {1, 2, 3, 4, 6, 8}) {
    const HostBVH b = build_bvh(g, leaf);
    CHECK(b.nnodes > 0 && b.tris.size() == 12 * (size_t)nf);
    CHECK(b.nodes.size() == 8 * (size_t)b.nnodes && b.wide.size() == 16 * (size_t)b.ninner);
    CHECK(b.wide4.size() == 32 * (size_t)b.ninner4);
  }
  {  // a mesh large enough for make_geometry to split over threads
    int64_t nv6 = 0, nf6 = 0;
    synth_sizes(6, &nv6, &nf6);
    std::vector<double> V6(3 * nv6);
    std::vector<int32_t> F6(3 * nf6);
    synth_ellipsoid(6, 13.0, 11.0, 9.0, 50, 0.05, 0.3, 0.2, 1.5, 7, V6.data(), F6.data());
    const HostGeometry g6 = make_geometry(V6.data(), nv6, F6.data(), nf6, nullptr, 1.0);
    check_geometry(g6, V6.data(), nv6, F6.data(), nullptr, 1.0);
  }
  {  // three bodies: one forest, one root per body
    HostGeometry gb = g;
    gb.body_off = {0, 100, 700, nf};
    const HostBVH b = build_bvh(gb, 4);
    CHECK(b.root_thr.size() == 3 && b.root_wide.size() == 3 && b.root_wide4.size() == 3);
  }

  // ---- OBJ reader: first shape, vertex normals, divisor 100 ---------------
  if (argc > 1) {
    std::vector<double> OV, ON;
    std::vector<int32_t> OF;
    read_obj(argv[1], OV, OF, ON);
    CHECK(!OF.empty() && OF.size() % 3 == 0);
    const int64_t onf = (int64_t)OF.size() / 3, onv = (int64_t)OV.size() / 3;
    const HostGeometry go = make_geometry(OV.data(), onv, OF.data(), onf,
                                          ON.empty() ? nullptr : ON.data(), 100.0);
    check_geometry(go, OV.data(), onv, OF.data(), ON.empty() ? nullptr : ON.data(), 100.0);
  }

  // ---- oracle --------------------------------------------------------------
  const float *tri = g.tri.data(), *cen = g.cen.data(), *nrm = g.nrm.data();
  // closest hits: BVH walk against brute force
  const int64_t nray = 4096;
  std::vector<float> orig(3 * nray), dir(3 * nray), t1(nray), t2(nray);
  std::vector<int32_t> h1(nray), h2(nray);
  for (int64_t k = 0; k < nray; ++k)
    for (int c = 0; c < 3; ++c) {
      orig[3 * k + c] = 0.1f * (float)c;
      dir[3 * k + c] = cen[3 * ((k * 7) % nf) + c];
    }
  CHECK(oracle_closest_hit(tri, nf, orig.data(), dir.data(), nray, 1, 4, h1.data(), t1.data()) == 0);
  CHECK(oracle_closest_hit(tri, nf, orig.data(), dir.data(), nray, 0, 4, h2.data(), t2.data()) == 0);
  CHECK(h1 == h2);
  // visibility (BVH and brute force) and form factors
  const int64_t words = (nf + 31) / 32;
  std::vector<uint32_t> bits(words * nf), bits2(words * nf);
  const int64_t nr = oracle_visibility(tri, cen, nrm, g.rad.data(), nf, 1e-5, 0, nf, 1, 4, bits.data());
  const int64_t nr2 = oracle_visibility(tri, cen, nrm, g.rad.data(), nf, 1e-5, 0, nf, 0, 4, bits2.data());
  CHECK(nr > 0 && nr == nr2 && bits == bits2);
  int64_t nb = 0;
  for (uint32_t w : bits) nb += __builtin_popcount(w);
  std::vector<int64_t> cp(nf + 1);
  std::vector<int32_t> ri(nb);
  std::vector<double> val(nb);
  const int64_t nnz = oracle_form_factors(cen, nrm, g.area.data(), nf, bits.data(), 0, nf, 0.3,
                                          nullptr, nb, cp.data(), ri.data(), val.data());
  CHECK(nnz > 0 && nnz <= nb && cp[nf] == nnz);
  // horizons over the body's own triangles: BVH walk against brute force
  const int nphi = 17;
  std::vector<double> hz(nphi * nf), hzo(nphi * nf);
  CHECK(oracle_horizons(cen, nrm, g.frame.data(), 0, nf, tri, nf, nphi, 1e-3, 1e-5, 1, 4, hz.data()) > 0);
  CHECK(oracle_horizons(cen, nrm, g.frame.data(), 0, nf, tri, nf, nphi, 1e-3, 1e-5, 0, 4, hzo.data()) > 0);
  CHECK(hz == hzo);
  // direct insolation for two suns over the fib spiral disk
  const int ndisk = 100;
  std::vector<double> disk(2 * ndisk);
  CHECK(oracle_fib_spiral(disk.data(), ndisk) == 0);
  const double sun[6] = {2.2739e8, 0.0, 1.0e7, -1.0e8, 2.0e8, 5.0e7};
  std::vector<double> E(2 * nf);
  CHECK(oracle_direct(cen, g.frame.data(), nf, hz.data(), nphi, sun, 2, disk.data(), ndisk, 4, E.data()) == 0);
  // Gauss-Seidel on the first sun
  std::vector<double> B(nf), res(50);
  const int sweeps = oracle_gauss_seidel(nf, cp.data(), ri.data(), val.data(), E.data(), 40, 1e-12,
                                         B.data(), res.data());
  CHECK(sweeps > 0 && sweeps <= 40);
  // conduction: one column, then a ragged two-grid run
  const int nz = 12;
  std::vector<double> z(nz), ti(nz, 200.0), rhoc(nz, 1.19177e6), T(nz, 233.0);
  for (int i = 0; i < nz; ++i) z[i] = 0.01 * (i + 1) * (i + 1);
  double Ts = 233.0, Fs = 0.0;
  CHECK(oracle_conductionQ(nz, z.data(), 60.0, 100.0, 120.0, T.data(), ti.data(), rhoc.data(), 0.95,
                           &Ts, 0.0, &Fs) == 0);
  CHECK(std::isfinite(Ts) && std::isfinite(Fs));
  const int nz2 = 7, nzmax = nz;
  std::vector<double> zz(z), tt(ti), rr(rhoc);
  zz.insert(zz.end(), z.begin(), z.begin() + nz2);
  tt.insert(tt.end(), ti.begin(), ti.begin() + nz2);
  rr.insert(rr.end(), rhoc.begin(), rhoc.begin() + nz2);
  const int64_t goff[3] = {0, nz, nz + nz2};
  const int64_t tf = 64;
  std::vector<int32_t> grid(tf);
  std::vector<double> scale(tf), Q(3 * tf), Qp(tf, 0.0), TT((nzmax + 1) * tf, 233.0), Fo(tf),
      Tser(3 * tf);
  for (int64_t f = 0; f < tf; ++f) {
    grid[f] = (int32_t)(f % 2);
    scale[f] = 0.5 + 0.01 * (double)f;
    for (int k = 0; k < 3; ++k) Q[k * tf + f] = 50.0 * (double)(k + 1);
  }
  CHECK(oracle_thermal_run(tf, goff, zz.data(), tt.data(), rr.data(), grid.data(), scale.data(), 0.95,
                           nzmax, 60.0, 3, Q.data(), Qp.data(), TT.data(), Fo.data(), Tser.data(), 4) == 0);
  // renderer ray-cast, both projections
  std::vector<int32_t> img(32 * 24);
  CHECK(oracle_render(tri, cen, nf, 0, 32, 24, 30.0, 20.0, 4, img.data()) == 0);
  CHECK(oracle_render(tri, cen, nf, 1, 32, 24, 0.0, 0.0, 4, img.data()) == 0);

  std::printf("host_check: %lld facets, %lld rays, nnz %lld, GS sweeps %d, %d failures\n",
              (long long)nf, (long long)nr, (long long)nnz, sweeps, fails);
  return fails ? 1 : 0;
}
