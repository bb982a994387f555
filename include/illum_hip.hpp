// illum_hip.hpp -- header-only C++ host API over the C ABI of illum_hip.h,
// keeping the reference's class and method names so host code written
// against illum_context / thermal_model carries over:
//
//   illum::illum_context      illum.hpp:6-48   (compute_F, make_horizons,
//                                               save/load_horizons,
//                                               get_direct_radiosity, get_num_faces,
//                                               set_horizon_obj_file)
//   illum::fib_spiral         illum.hpp:50
//   illum::forward_solve-style radiosity      illum_cli.cpp:425-440
//   illum::thermal_model      thermal.hpp:7-105 (step, get_radiosity, T)
//   illum::conductionQ        conduction.hpp:16-28
//
// Matrices are plain column-major std::vector<double> with explicit sizes
// (Armadillo is not a dependency); the sparse form-factor matrix stays on the
// GPU inside illum::sp_mat and can be copied out as CSC (the reference's
// arma::sp_mat layout) or CSR.  Errors throw illum::error carrying the status.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "illum_hip.h"

namespace illum {

struct error : std::runtime_error {
  int status;
  error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

inline void check(int s) {
  if (s != ILLUM_OK) throw error(s, illum_last_error());
}

// Owning wrapper of one C handle type.
template <class H, int (*Destroy)(H *)>
struct handle {
  H *h = nullptr;
  handle() = default;
  explicit handle(H *p) : h(p) {}
  handle(const handle &) = delete;
  handle &operator=(const handle &) = delete;
  handle(handle &&o) noexcept : h(o.h) { o.h = nullptr; }
  handle &operator=(handle &&o) noexcept {
    if (this != &o) { reset(); h = o.h; o.h = nullptr; }
    return *this;
  }
  ~handle() { reset(); }
  void reset() { if (h) Destroy(h); h = nullptr; }
  H *get() const { return h; }
};

using ctx_handle = handle<illum_ctx, illum_ctx_destroy>;

class context {
 public:
  explicit context(int device = 0) {
    illum_ctx *c = nullptr;
    check(illum_ctx_create(device, &c));
    h_ = ctx_handle(c);
  }
  illum_ctx *get() const { return h_.get(); }
  void synchronize() const { check(illum_ctx_synchronize(h_.get())); }

 private:
  ctx_handle h_;
};

// Sparse N x N form-factor matrix resident on the GPU.
class sp_mat {
 public:
  sp_mat() = default;
  explicit sp_mat(illum_sparse *s) : h_(s) { check(illum_sparse_info(s, &n_, &nnz_)); }
  int64_t n_rows() const { return n_; }
  int64_t n_cols() const { return n_; }
  int64_t n_nonzero() const { return nnz_; }
  illum_sparse *get() const { return h_.get(); }
  // CSC arrays (column j = source facet), as arma::sp_mat stores them.
  void csc(std::vector<int64_t> &colptr, std::vector<int32_t> &rowidx,
           std::vector<double> &val) const {
    colptr.resize(n_ + 1); rowidx.resize(nnz_); val.resize(nnz_);
    check(illum_sparse_copy_csc(h_.get(), colptr.data(), rowidx.data(), val.data()));
  }
  void csr(std::vector<int64_t> &rowptr, std::vector<int32_t> &colidx,
           std::vector<double> &val) const {
    rowptr.resize(n_ + 1); colidx.resize(nnz_); val.resize(nnz_);
    check(illum_sparse_copy_csr(h_.get(), rowptr.data(), colidx.data(), val.data()));
  }
  static sp_mat from_csr(const context &ctx, int64_t n, const std::vector<int64_t> &rowptr,
                         const std::vector<int32_t> &colidx, const std::vector<double> &val) {
    illum_sparse *s = nullptr;
    check(illum_sparse_create_csr(ctx.get(), n, rowptr.back(), rowptr.data(), colidx.data(),
                                  val.data(), &s));
    return sp_mat(s);
  }

 private:
  handle<illum_sparse, illum_sparse_destroy> h_;
  int64_t n_ = 0, nnz_ = 0;
};

inline std::vector<double> fib_spiral(int n) {  // n x 2 column-major
  std::vector<double> xy(2 * (size_t)n);
  check(illum_fib_spiral(xy.data(), n));
  return xy;
}

class illum_context {
 public:
  // From an OBJ file with the reference loader's semantics (vertices / 100).
  illum_context(const context &ctx, const std::string &obj_path, int leaf_size = 4,
                double vertex_divisor = 100.0)
      : ctx_(&ctx) {
    illum_mesh *m = nullptr;
    check(illum_mesh_from_obj(ctx.get(), obj_path.c_str(), vertex_divisor, &m));
    init(m, leaf_size);
  }
  // From vertex (nv x 3) / face (nf x 3) arrays, row-major.
  illum_context(const context &ctx, const std::vector<double> &V, const std::vector<int32_t> &F,
                int leaf_size = 4, double vertex_divisor = 1.0)
      : ctx_(&ctx) {
    illum_mesh *m = nullptr;
    check(illum_mesh_create(ctx.get(), V.data(), (int64_t)V.size() / 3, F.data(),
                            (int64_t)F.size() / 3, nullptr, vertex_divisor, &m));
    init(m, leaf_size);
  }

  int get_num_faces() const { return (int)num_faces; }

  void set_horizon_obj_file(const std::string &path, double vertex_divisor = 100.0) {
    illum_mesh *m = nullptr;
    check(illum_mesh_from_obj(ctx_->get(), path.c_str(), vertex_divisor, &m));
    horizon_mesh_ = mesh_handle(m);
    illum_bvh *b = nullptr;
    check(illum_bvh_build(m, leaf_size_, &b));
    horizon_bvh_ = bvh_handle(b);
  }

  // albedo: scalar (F *= rho) or per-facet vector (F.row(i) *= rho(i)), illum.cpp:225-234.
  sp_mat compute_F(double albedo, double offset = 1e-5) {
    illum_sparse *s = nullptr;
    check(illum_form_factors(bvh_.get(), albedo, nullptr, offset, &s));
    return sp_mat(s);
  }
  sp_mat compute_F(const std::vector<double> &albedo, double offset = 1e-5) {
    if ((int64_t)albedo.size() != num_faces) throw error(ILLUM_ERR_ARG, "albedo size");
    illum_sparse *s = nullptr;
    check(illum_form_factors(bvh_.get(), 0.0, albedo.data(), offset, &s));
    return sp_mat(s);
  }

  void make_horizons(int nphi = 361, double theta_eps = 1e-3, double offset = 1e-5) {
    horizons.assign((size_t)nphi * num_faces, 0.0);
    nphi_ = nphi;
    check(illum_horizons(bvh_.get(), horizon_bvh_.get(), nphi, theta_eps, offset,
                         horizons.data(), 0));
  }

  // sun_position: 3 x S column-major (km); returns N x S column-major.
  std::vector<double> get_direct_radiosity(const std::vector<double> &sun_position,
                                           const std::vector<double> &disk_xy) const {
    const int S = (int)(sun_position.size() / 3);
    const int ndisk = (int)(disk_xy.size() / 2);
    std::vector<double> E((size_t)num_faces * S);
    check(illum_direct_radiosity(mesh_.get(), horizons.data(), nphi_, sun_position.data(), S,
                                 disk_xy.data(), ndisk, E.data(), 0));
    return E;
  }

  // A horizon map the caller already holds: nphi x N column-major.
  void set_horizons(std::vector<double> h, int nphi) {
    if ((int64_t)h.size() != (int64_t)nphi * num_faces) throw error(ILLUM_ERR_SHAPE, "horizon map size");
    horizons = std::move(h);
    nphi_ = nphi;
  }

  // save_horizons / load_horizons (illum.hpp:25-27): the nphi x N map in
  // Armadillo's dense binary format, "ARMA_MAT_BIN_FN008\n<rows> <cols>\n"
  // then the FP64 values column by column (what arma::mat::save writes).
  void save_horizons(const std::string &path) const {
    std::FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw error(ILLUM_ERR_IO, "cannot write " + path);
    const std::string head = "ARMA_MAT_BIN_FN008\n" + std::to_string(nphi_) + " " +
                             std::to_string(num_faces) + "\n";
    const bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size() &&
                    std::fwrite(horizons.data(), 8, horizons.size(), f) == horizons.size();
    std::fclose(f);
    if (!ok) throw error(ILLUM_ERR_IO, "short write to " + path);
  }
  void load_horizons(const std::string &path) {
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw error(ILLUM_ERR_IO, "cannot read " + path);
    char tag[20] = {0};
    long long rows = 0, cols = 0;
    const bool head_ok = std::fread(tag, 1, 19, f) == 19 &&
                         std::string(tag, 19) == "ARMA_MAT_BIN_FN008\n" &&
                         std::fscanf(f, "%lld %lld", &rows, &cols) == 2 && std::fgetc(f) == '\n';
    if (!head_ok || rows < 2 || cols != num_faces) {
      std::fclose(f);
      throw error(head_ok ? ILLUM_ERR_SHAPE : ILLUM_ERR_IO,
                  path + ": not an nphi x N FP64 Armadillo matrix");
    }
    std::vector<double> h((size_t)rows * (size_t)cols);
    const bool ok = std::fread(h.data(), 8, h.size(), f) == h.size();
    std::fclose(f);
    if (!ok) throw error(ILLUM_ERR_IO, path + ": truncated");
    set_horizons(std::move(h), (int)rows);
  }

  // The renderer's ray-cast (illum_renderer_cli.cpp:174-261): hit facet per
  // pixel (-1 = none), height x width row-major; mode 0 orthographic, 1
  // equal-area cylindrical.  With data (N values), value gets data[facet]
  // clamped to [minval, maxval] (NaN where nothing is hit).
  std::vector<int32_t> render(int mode, int width, int height, double az, double el,
                              const std::vector<double> *data = nullptr,
                              std::vector<double> *value = nullptr,
                              double minval = -HUGE_VAL, double maxval = HUGE_VAL) const {
    std::vector<int32_t> facet((size_t)width * height);
    if (value) value->assign((size_t)width * height, 0.0);
    if (data && (int64_t)data->size() != num_faces) throw error(ILLUM_ERR_SHAPE, "data size");
    check(illum_render(bvh_.get(), mode, width, height, az, el, data ? data->data() : nullptr,
                       minval, maxval, facet.data(), value && data ? value->data() : nullptr,
                       0));
    return facet;
  }

  int nphi() const { return nphi_; }
  illum_bvh *bvh() const { return bvh_.get(); }

  int64_t num_faces = 0;
  std::vector<double> horizons;  // nphi x N column-major (illum.hpp:47)

 private:
  using mesh_handle = handle<illum_mesh, illum_mesh_destroy>;
  using bvh_handle = handle<illum_bvh, illum_bvh_destroy>;
  void init(illum_mesh *m, int leaf_size) {
    mesh_ = mesh_handle(m);
    check(illum_mesh_num_faces(m, &num_faces));
    illum_bvh *b = nullptr;
    check(illum_bvh_build(m, leaf_size, &b));
    bvh_ = bvh_handle(b);
    leaf_size_ = leaf_size;
  }
  const context *ctx_;
  mesh_handle mesh_, horizon_mesh_;
  bvh_handle bvh_, horizon_bvh_;
  int leaf_size_ = 4;
  int nphi_ = 0;
};

// Solves (I - F) B = E for S right-hand sides (N x S column-major).
struct radiosity_result {
  std::vector<double> B;
  std::vector<double> residuals;  // relative residual after each FP64 check
  int products = 0;
};
inline radiosity_result solve_radiosity(const sp_mat &F, const std::vector<double> &E, int S,
                                        int method = ILLUM_RAD_MIXED, double rel_tol = 1e-8,
                                        int max_iter = 200) {
  radiosity_result r;
  r.B.assign(E.size(), 0.0);
  r.residuals.assign((size_t)max_iter + 1, 0.0);
  check(illum_radiosity_solve(F.get(), E.data(), S, method, rel_tol, max_iter, r.B.data(),
                              r.residuals.data(), &r.products, 0));
  return r;
}

class thermal_model {
 public:
  // thermal.hpp:17-59: one depth grid for all facets, scalar thermal inertia,
  // scalar or per-layer rhoc, scalar initial temperature.
  thermal_model(const context &ctx, int nfaces, const std::vector<double> &depths,
                double thermal_inertia, const std::vector<double> &rhoc, double init_temp,
                double emissivity = 0.999, int precision = ILLUM_THERM_FP64)
      : nz((int)depths.size()), nf_(nfaces) {
    std::vector<double> ti(depths.size(), thermal_inertia);
    std::vector<double> rc = rhoc.size() == 1 ? std::vector<double>(depths.size(), rhoc[0]) : rhoc;
    if (rc.size() != depths.size()) throw error(ILLUM_ERR_ARG, "rhoc size");
    const int64_t off[2] = {0, (int64_t)depths.size()};
    illum_thermal *t = nullptr;
    check(illum_thermal_create(ctx.get(), nfaces, 1, off, depths.data(), ti.data(), rc.data(),
                               nullptr, nullptr, emissivity, init_temp, nullptr, precision, &t));
    h_ = th_handle(t);
  }
  void step(double dt, const std::vector<double> &Q) {
    if ((int64_t)Q.size() != nf_) throw error(ILLUM_ERR_SHAPE, "Q size");
    check(illum_thermal_step(h_.get(), dt, Q.data(), 0));
    t += dt;
  }
  // -F (thermal.hpp:61-63)
  std::vector<double> get_radiosity() const {
    std::vector<double> F(nf_);
    check(illum_thermal_get(h_.get(), nullptr, nullptr, F.data(), 0));
    for (auto &x : F) x = -x;
    return F;
  }
  // (nz+1) x N column-major, row 0 = surface.
  std::vector<double> T() const {
    std::vector<double> out((size_t)(nz + 1) * nf_);
    check(illum_thermal_get(h_.get(), out.data(), nullptr, nullptr, 0));
    return out;
  }
  int nz;
  double t = 0;

 private:
  using th_handle = handle<illum_thermal, illum_thermal_destroy>;
  th_handle h_;
  int64_t nf_;
};

inline void conductionQ(int nz, const double *z, double dt, double Qn, double Qnp1, double *T,
                        const double *ti, const double *rhoc, double emiss, double *Tsurf,
                        double Fgeotherm, double *Fsurf) {
  check(illum_conductionQ(nz, z, dt, Qn, Qnp1, T, ti, rhoc, emiss, Tsurf, Fgeotherm, Fsurf));
}

}  // namespace illum
