// Renderer ray-cast (illum_renderer_cli.cpp:174-261): the facet hit by each
// pixel's ray and the per-facet field value there.
//
// One lane per pixel.  Camera set-up follows the reference: the model centre
// is the mean of the facet centroids (FP32 sums in facet order), the model
// radius the largest centroid distance, and the camera sits at 1.1 x that
// radius.
//   mode 0 (orthographic): a camera at azimuth az, elevation el (degrees)
//     looking at the centre; pixel (i, j) casts along the view direction from
//     p_cam + y_eye n_up + x_eye n_right.
//   mode 1 (equal-area cylindrical): pixel (i, j) looks at the centre from
//     sph2xyz(2 pi j / W - az, asin(2 (i + 1) / (H + 1) - 1) - el); az and el
//     are used as given (the reference converts degrees only in the
//     orthographic branch).
// Vector arithmetic is FP32 component-wise; scalars are FP64 rounded to FP32
// where they meet a vector; |v| = sqrt of (x*x + y*y) + z*z, taken in FP64 of
// the FP32 sum and rounded (the correctly rounded FP32 square root).  The
// per-row and per-column trigonometry of mode 1 is tabulated on the host so
// directions do not depend on the device's transcendental functions.
// Closest hit is taken over every body's tree (as illum_bvh_closest_hit).
#include "hip_util.hpp"
#include "ray.hpp"

#include <cmath>
#include <vector>

namespace illum {

namespace {

struct Cam {
  float pc[3], up[3], right[3], eye[3], c[3];
  double rad, aspect;
};

__device__ __forceinline__ void unit_neg(float x, float y, float z, float *d) {
  const float l = (float)sqrt((double)((x * x + y * y) + z * z));
  d[0] = -(x / l);
  d[1] = -(y / l);
  d[2] = -(z / l);
}

constexpr int RENDER_STACK = 8;

__global__ __launch_bounds__(256) void render_kernel(
    int mode, int width, int height, Cam cam, const double *__restrict__ row_tab,
    const double *__restrict__ col_tab, const float4 *__restrict__ nodes,
    const float4 *__restrict__ wide, const float4 *__restrict__ tris,
    const int32_t *__restrict__ roots, int nbody, const double *__restrict__ data,
    double lo, double hi, int32_t *__restrict__ facet, double *__restrict__ value) {
  __shared__ int stack[RENDER_STACK * 256];
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)width * height) return;
  const int i = (int)(k / width), j = (int)(k - (int64_t)i * width);
  float o[3], d[3];
  if (mode == 0) {
    double y_eye = 2 * (1. - (double)i / (height - 1)) - 1;
    y_eye *= cam.rad;
    double x_eye = 2 * (double)j / (width - 1) - 1;
    x_eye *= cam.aspect * cam.rad;
    const float ys = (float)y_eye, xs = (float)x_eye;
    for (int a = 0; a < 3; ++a) {
      o[a] = (cam.pc[a] + ys * cam.up[a]) + xs * cam.right[a];
      d[a] = cam.eye[a];
    }
  } else {
    // row_tab: sin, cos of (theta_i - el); col_tab: cos, sin of (phi_j - az)
    const double st = row_tab[2 * i], ct = row_tab[2 * i + 1];
    const double cp = col_tab[2 * j], sp = col_tab[2 * j + 1];
    const float px = (float)((cam.rad * cp) * st), py = (float)((cam.rad * sp) * st),
                pz = (float)(cam.rad * ct);
    unit_neg(px, py, pz, d);
    o[0] = px + cam.c[0];
    o[1] = py + cam.c[1];
    o[2] = pz + cam.c[2];
  }
  const DRay r = make_ray(o[0], o[1], o[2], d[0], d[1], d[2]);
  unsigned nbox = 0, ntri = 0;
  int best = -1;
  float tbest = __builtin_huge_valf();
  for (int b = 0; b < nbody; ++b) {
    float t;
    int h = -2;
    if (wide)
      h = trace_closest_wide(wide, tris, r, roots[4 * b + 1], __builtin_huge_valf(),
                             stack + threadIdx.x, 256, RENDER_STACK, &t, nbox, ntri);
    if (h == -2) h = trace_closest(nodes, tris, r, roots[4 * b], &t, nbox, ntri);
    if (h >= 0 && (t < tbest || (t == tbest && h < best))) {
      tbest = t;
      best = h;
    }
  }
  if (facet) facet[k] = best;
  if (value) {
    double v = __builtin_nan("");
    if (best >= 0 && data) v = fmin(hi, fmax(lo, data[best]));
    value[k] = v;
  }
}

}  // namespace

void render(illum_bvh *bvh, int mode, int width, int height, double az, double el,
            const double *data, double lo, double hi, int32_t *facet, double *value,
            int flags) {
  illum_mesh *mesh = bvh->mesh;
  illum_ctx *ctx = mesh->ctx;
  hipStream_t s = stream_of(ctx);
  check_bvh_bodies(bvh);
  if (mode != 0 && mode != 1) throw Error(ILLUM_ERR_ARG, "mode must be 0 (orthographic) or 1");
  if (width < 2 || height < 2) throw Error(ILLUM_ERR_ARG, "image must be at least 2 x 2");
  if (value && !data) throw Error(ILLUM_ERR_ARG, "values need per-facet data");
  const int64_t nf = mesh->g.nf;
  const double PI = 3.141592653589793;
  const float *cen = mesh->g.cen.data();
  // Model centre and radius (illum_renderer_cli.cpp:174-184).
  float c[3] = {0.0f, 0.0f, 0.0f};
  for (int64_t f = 0; f < nf; ++f)
    for (int a = 0; a < 3; ++a) c[a] = c[a] + cen[3 * f + a];
  for (int a = 0; a < 3; ++a) c[a] = c[a] / (float)nf;
  double r_model = 0.0;
  for (int64_t f = 0; f < nf; ++f) {
    const float x = cen[3 * f] - c[0], y = cen[3 * f + 1] - c[1], z = cen[3 * f + 2] - c[2];
    r_model = std::fmax(r_model, (double)(float)std::sqrt((double)((x * x + y * y) + z * z)));
  }
  Cam cam{};
  cam.rad = 1.1 * r_model;
  cam.aspect = (double)width / height;
  for (int a = 0; a < 3; ++a) cam.c[a] = c[a];
  auto sph = [](double phi, double theta, double rad, float *o) {
    o[0] = (float)((rad * std::cos(phi)) * std::sin(theta));
    o[1] = (float)((rad * std::sin(phi)) * std::sin(theta));
    o[2] = (float)(rad * std::cos(theta));
  };
  std::vector<double> rows, cols;
  if (mode == 0) {
    const double azr = az * (PI / 180), elr = el * (PI / 180);
    float p[3];
    sph(azr, elr, cam.rad, p);
    const float l = (float)std::sqrt((double)((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]));
    for (int a = 0; a < 3; ++a) cam.eye[a] = -(p[a] / l);
    sph(azr, elr - PI / 2, 1.0, cam.up);
    const float *e = cam.eye, *u = cam.up;
    cam.right[0] = e[1] * u[2] - e[2] * u[1];
    cam.right[1] = e[2] * u[0] - e[0] * u[2];
    cam.right[2] = e[0] * u[1] - e[1] * u[0];
    for (int a = 0; a < 3; ++a) cam.pc[a] = p[a] + c[a];
  } else {
    rows.resize(2 * (size_t)height);
    cols.resize(2 * (size_t)width);
    for (int i = 0; i < height; ++i) {
      const double theta = std::asin(2 * (double)(i + 1) / (height + 1) - 1);
      rows[2 * i] = std::sin(theta - el);
      rows[2 * i + 1] = std::cos(theta - el);
    }
    for (int j = 0; j < width; ++j) {
      const double phi = 2 * PI * (double)j / width;
      cols[2 * j] = std::cos(phi - az);
      cols[2 * j + 1] = std::sin(phi - az);
    }
  }
  const int64_t npix = (int64_t)width * height;
  DBuf<double> d_rows(rows.empty() ? 1 : rows.size()), d_cols(cols.empty() ? 1 : cols.size());
  if (!rows.empty()) {
    ILLUM_HIP(hipMemcpyAsync(d_rows.p, rows.data(), 8 * rows.size(), hipMemcpyHostToDevice, s));
    ILLUM_HIP(hipMemcpyAsync(d_cols.p, cols.data(), 8 * cols.size(), hipMemcpyHostToDevice, s));
  }
  const bool in_dev = flags & ILLUM_IN_ON_DEVICE, out_dev = flags & ILLUM_OUT_ON_DEVICE;
  DBuf<double> d_data;
  const double *dd = data;
  if (data && !in_dev) {
    d_data.reset(nf);
    ILLUM_HIP(hipMemcpyAsync(d_data.p, data, 8 * nf, hipMemcpyHostToDevice, s));
    dd = d_data.p;
  }
  DBuf<int32_t> d_facet;
  DBuf<double> d_value;
  int32_t *fo = facet;
  double *vo = value;
  if (facet && !out_dev) { d_facet.reset(npix); fo = d_facet.p; }
  if (value && !out_dev) { d_value.reset(npix); vo = d_value.p; }
  {
    StageTimer st(ctx, "render");
    render_kernel<<<(unsigned)((npix + 255) / 256), 256, 0, s>>>(
        mode, width, height, cam, d_rows.p, d_cols.p, (const float4 *)bvh->d_nodes,
        (const float4 *)bvh->d_wide, (const float4 *)bvh->d_tris, bvh->d_roots, bvh->nbody, dd,
        lo, hi, fo, vo);
    ILLUM_LAUNCH_CHECK();
  }
  if (facet && !out_dev)
    ILLUM_HIP(hipMemcpyAsync(facet, fo, 4 * npix, hipMemcpyDeviceToHost, s));
  if (value && !out_dev)
    ILLUM_HIP(hipMemcpyAsync(value, vo, 8 * npix, hipMemcpyDeviceToHost, s));
  sync(ctx);
}

}  // namespace illum
