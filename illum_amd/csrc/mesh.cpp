// Mesh ingestion: seeded synthetic cratered ellipsoids, a minimal OBJ reader,
// and the per-facet geometry every kernel consumes.
//
// Reference behaviour followed:
//   obj_util.cpp:43-93  vertices as FP32, face normal = normalize((v1-v0)x(v2-v0))
//                       flipped to agree with the vertex normals indexed by
//                       vertex index, vertices divided by 100 (FP32)
//   illum.cpp:244-262   Frenet frame columns [b t n], t = normalise(v1-v0),
//                       b = t x n
//   illum.cpp:101-105   facet area |(v1-v0)x(v2-v0)|/2
// Compiled with -ffp-contract=off: every FP32/FP64 expression below is
// evaluated exactly as written (IEEE round-to-nearest, no fused ops).
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <thread>
#include <unordered_map>

namespace illum {

// ---------------------------------------------------------------------------
// SplitMix64: the generator's only source of randomness (fully specified, so
// every host produces identical meshes for a seed).
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

void synth_sizes(int subdiv, int64_t *nv, int64_t *nf) {
  if (subdiv < 0 || subdiv > 9) throw Error(ILLUM_ERR_ARG, "subdiv must be in [0, 9]");
  int64_t p = 1;
  for (int k = 0; k < subdiv; ++k) p *= 4;
  if (nv) *nv = 10 * p + 2;
  if (nf) *nf = 20 * p;
}

void synth_ellipsoid(int subdiv, double a, double b, double c, int ncraters, double dmin,
                     double dmax, double wmin, double wmax, uint64_t seed, double *V,
                     int32_t *F) {
  int64_t nv_expect, nf_expect;
  synth_sizes(subdiv, &nv_expect, &nf_expect);
  if (!(a > 0 && b > 0 && c > 0)) throw Error(ILLUM_ERR_ARG, "semi-axes must be positive");
  if (ncraters < 0) throw Error(ILLUM_ERR_ARG, "ncraters must be >= 0");

  // Icosahedron on the unit sphere.
  const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<double> P = {
      -1, phi, 0,  1, phi, 0,  -1, -phi, 0,  1, -phi, 0,
      0, -1, phi,  0, 1, phi,  0, -1, -phi,  0, 1, -phi,
      phi, 0, -1,  phi, 0, 1,  -phi, 0, -1,  -phi, 0, 1};
  for (int v = 0; v < 12; ++v) {
    double *x = &P[3 * v];
    double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    x[0] /= r; x[1] /= r; x[2] /= r;
  }
  std::vector<int32_t> T = {0, 11, 5,  0, 5, 1,   0, 1, 7,   0, 7, 10,  0, 10, 11,
                            1, 5, 9,   5, 11, 4,  11, 10, 2, 10, 7, 6,  7, 1, 8,
                            3, 9, 4,   3, 4, 2,   3, 2, 6,   3, 6, 8,   3, 8, 9,
                            4, 9, 5,   2, 4, 11,  6, 2, 10,  8, 6, 7,   9, 8, 1};
  // Make every face wind outward.
  for (int f = 0; f < 20; ++f) {
    const double *p0 = &P[3 * T[3 * f]], *p1 = &P[3 * T[3 * f + 1]], *p2 = &P[3 * T[3 * f + 2]];
    double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                   e1[0] * e2[1] - e1[1] * e2[0]};
    double s = n[0] * (p0[0] + p1[0] + p2[0]) + n[1] * (p0[1] + p1[1] + p2[1]) +
               n[2] * (p0[2] + p1[2] + p2[2]);
    if (s < 0) std::swap(T[3 * f + 1], T[3 * f + 2]);
  }

  // Loop subdivision of the connectivity; midpoints projected to the sphere.
  // Child faces of face f are 4f..4f+3, which keeps facet order spatially
  // coherent (neighbouring indices are neighbouring facets).
  for (int level = 0; level < subdiv; ++level) {
    std::unordered_map<uint64_t, int32_t> mid;
    mid.reserve(T.size());
    auto midpoint = [&](int32_t i, int32_t j) -> int32_t {
      uint64_t key = i < j ? ((uint64_t)i << 32) | (uint32_t)j : ((uint64_t)j << 32) | (uint32_t)i;
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      double m[3] = {0.5 * (P[3 * i] + P[3 * j]), 0.5 * (P[3 * i + 1] + P[3 * j + 1]),
                     0.5 * (P[3 * i + 2] + P[3 * j + 2])};
      double r = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
      int32_t id = (int32_t)(P.size() / 3);
      P.push_back(m[0] / r); P.push_back(m[1] / r); P.push_back(m[2] / r);
      mid.emplace(key, id);
      return id;
    };
    std::vector<int32_t> T2;
    T2.reserve(T.size() * 4);
    for (size_t f = 0; f < T.size() / 3; ++f) {
      int32_t v0 = T[3 * f], v1 = T[3 * f + 1], v2 = T[3 * f + 2];
      int32_t m01 = midpoint(v0, v1), m12 = midpoint(v1, v2), m20 = midpoint(v2, v0);
      int32_t kids[12] = {v0, m01, m20, v1, m12, m01, v2, m20, m12, m01, m12, m20};
      T2.insert(T2.end(), kids, kids + 12);
    }
    T.swap(T2);
  }
  const int64_t nv = (int64_t)(P.size() / 3), nf = (int64_t)(T.size() / 3);
  if (nv != nv_expect || nf != nf_expect) throw Error(ILLUM_ERR_ARG, "internal: icosphere size");

  // Ellipsoid.
  for (int64_t v = 0; v < nv; ++v) {
    V[3 * v] = a * P[3 * v];
    V[3 * v + 1] = b * P[3 * v + 1];
    V[3 * v + 2] = c * P[3 * v + 2];
  }

  // Craters: centre direction uniform on the sphere, mapped onto the
  // undisplaced ellipsoid; depth and width uniform; Gaussian radial dent
  // depth * exp(-s^2 / (2 w^2)), s = chord distance to the centre point.
  SplitMix64 rng(seed);
  std::vector<double> cx(ncraters), cy(ncraters), cz(ncraters), dep(ncraters), wid(ncraters);
  for (int k = 0; k < ncraters; ++k) {
    double zc = 2.0 * rng.uniform() - 1.0;
    double az = 2.0 * M_PI * rng.uniform();
    double rr = std::sqrt(std::max(0.0, 1.0 - zc * zc));
    cx[k] = a * rr * std::cos(az);
    cy[k] = b * rr * std::sin(az);
    cz[k] = c * zc;
    dep[k] = dmin + (dmax - dmin) * rng.uniform();
    wid[k] = wmin + (wmax - wmin) * rng.uniform();
  }
  if (ncraters > 0) {
    for (int64_t v = 0; v < nv; ++v) {
      double x = V[3 * v], y = V[3 * v + 1], z = V[3 * v + 2];
      double dent = 0.0;
      for (int k = 0; k < ncraters; ++k) {
        double dx = x - cx[k], dy = y - cy[k], dz = z - cz[k];
        double s2 = dx * dx + dy * dy + dz * dz;
        dent += dep[k] * std::exp(-s2 / (2.0 * wid[k] * wid[k]));
      }
      double r = std::sqrt(x * x + y * y + z * z);
      V[3 * v] = x - dent * x / r;
      V[3 * v + 1] = y - dent * y / r;
      V[3 * v + 2] = z - dent * z / r;
    }
  }
  std::memcpy(F, T.data(), sizeof(int32_t) * T.size());
}

// ---------------------------------------------------------------------------
void read_obj(const char *path, std::vector<double> &V, std::vector<int32_t> &F,
              std::vector<double> &VN) {
  std::ifstream in(path);
  if (!in) throw Error(ILLUM_ERR_IO, std::string("cannot open OBJ file ") + path);
  V.clear(); F.clear(); VN.clear();
  std::string line;
  int64_t lineno = 0;
  // The reference keeps only the first shape (obj_util.cpp:34, shape_index
  // 0): the faces up to the first "g" or "o" line that follows a face.
  // Vertices and normals form one pool for every shape, so they are read to
  // the end of the file; later faces are skipped.
  bool first_shape_done = false;
  while (std::getline(in, line)) {
    ++lineno;
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    std::string tag;
    ss >> tag;
    if (tag == "g" || tag == "o") {
      if (!F.empty()) first_shape_done = true;
    } else if (tag == "v") {
      double x, y, z;
      if (!(ss >> x >> y >> z)) throw Error(ILLUM_ERR_IO, "bad vertex at line " + std::to_string(lineno));
      V.push_back(x); V.push_back(y); V.push_back(z);
    } else if (tag == "vn") {
      double x, y, z;
      if (!(ss >> x >> y >> z)) throw Error(ILLUM_ERR_IO, "bad normal at line " + std::to_string(lineno));
      VN.push_back(x); VN.push_back(y); VN.push_back(z);
    } else if (tag == "f" && !first_shape_done) {
      std::vector<int32_t> idx;
      std::string tok;
      const int64_t nvert = (int64_t)(V.size() / 3);
      while (ss >> tok) {
        long k = std::strtol(tok.c_str(), nullptr, 10);  // vertex index before any '/'
        if (k == 0) throw Error(ILLUM_ERR_IO, "bad face index at line " + std::to_string(lineno));
        // negative: relative to the vertices read so far; positive indices
        // are checked once every vertex is read
        long v = k > 0 ? k - 1 : nvert + k;
        if (v < 0) throw Error(ILLUM_ERR_IO, "face index out of range at line " + std::to_string(lineno));
        idx.push_back((int32_t)v);
      }
      if (idx.size() < 3) throw Error(ILLUM_ERR_IO, "face with < 3 vertices at line " + std::to_string(lineno));
      for (size_t t = 1; t + 1 < idx.size(); ++t) {  // fan triangulation
        F.push_back(idx[0]); F.push_back(idx[t]); F.push_back(idx[t + 1]);
      }
    }
  }
  if (F.empty()) throw Error(ILLUM_ERR_IO, std::string("no faces in OBJ file ") + path);
  const int64_t nvert = (int64_t)(V.size() / 3);
  for (int32_t v : F)
    if (v >= nvert) throw Error(ILLUM_ERR_IO, std::string("face index out of range in ") + path);
}

// ---------------------------------------------------------------------------
HostGeometry make_geometry(const double *V, int64_t nv, const int32_t *F, int64_t nf,
                           const double *VN, double vertex_divisor) {
  if (!V || !F || nv <= 0 || nf <= 0) throw Error(ILLUM_ERR_ARG, "mesh needs vertices and faces");
  if (nf >= (1 << 24)) throw Error(ILLUM_ERR_CAPACITY, "at most 16,777,215 facets are supported");
  if (!(vertex_divisor > 0)) throw Error(ILLUM_ERR_ARG, "vertex_divisor must be positive");
  const float div = (float)vertex_divisor;
  HostGeometry g;
  g.nf = nf;
  g.tri.resize(9 * nf);
  g.cen.resize(3 * nf);
  g.nrm.resize(3 * nf);
  g.rad.resize(nf);
  g.area.resize(nf);
  g.frame.resize(9 * nf);
  for (int64_t k = 0; k < 3 * nf; ++k)
    if (F[k] < 0 || F[k] >= nv) throw Error(ILLUM_ERR_ARG, "face index out of range");
  // Every facet's records depend on its own vertices only: facet ranges are
  // computed on all host threads (the same values as one thread).
  auto facets = [&](int64_t f0, int64_t f1) {
  for (int64_t f = f0; f < f1; ++f) {
    // u: the FP32 vertices as read; v: divided by the vertex divisor.  The
    // face normal comes from u and everything else from v, as the reference
    // computes n before building Tri(v0/100., v1/100., v2/100., n)
    // (obj_util.cpp:70, :92).
    float u[3][3], v[3][3];
    for (int k = 0; k < 3; ++k) {
      int32_t id = F[3 * f + k];
      for (int c = 0; c < 3; ++c) {
        u[k][c] = (float)V[3 * id + c];
        v[k][c] = vertex_divisor == 1.0 ? u[k][c] : u[k][c] / div;
        g.tri[9 * f + 3 * k + c] = v[k][c];
      }
    }
    float e1[3] = {u[1][0] - u[0][0], u[1][1] - u[0][1], u[1][2] - u[0][2]};
    float e2[3] = {u[2][0] - u[0][0], u[2][1] - u[0][1], u[2][2] - u[0][2]};
    float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                   e1[0] * e2[1] - e1[1] * e2[0]};
    float len = std::sqrt((cr[0] * cr[0] + cr[1] * cr[1]) + cr[2] * cr[2]);
    float n[3] = {cr[0] / len, cr[1] / len, cr[2] / len};
    if (VN) {
      float s[3];
      for (int c = 0; c < 3; ++c) {
        float n0 = (float)VN[3 * F[3 * f] + c], n1 = (float)VN[3 * F[3 * f + 1] + c],
              n2 = (float)VN[3 * F[3 * f + 2] + c];
        s[c] = (n0 + n1) + n2;
      }
      if ((n[0] * s[0] + n[1] * s[1]) + n[2] * s[2] < 0) {
        n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2];
      }
    }
    float p[3];
    for (int c = 0; c < 3; ++c) p[c] = ((v[0][c] + v[1][c]) + v[2][c]) / 3.0f;
    float r = 0.0f;
    for (int k = 0; k < 3; ++k) {
      float dx = v[k][0] - p[0], dy = v[k][1] - p[1], dz = v[k][2] - p[2];
      float d = std::sqrt((dx * dx + dy * dy) + dz * dz);
      r = d > r ? d : r;
    }
    for (int c = 0; c < 3; ++c) {
      g.cen[3 * f + c] = p[c];
      g.nrm[3 * f + c] = n[c];
    }
    g.rad[f] = r;
    // FP64 area from the FP32 vertices.
    double d1[3], d2[3];
    for (int c = 0; c < 3; ++c) {
      d1[c] = (double)v[1][c] - (double)v[0][c];
      d2[c] = (double)v[2][c] - (double)v[0][c];
    }
    double cd[3] = {d1[1] * d2[2] - d1[2] * d2[1], d1[2] * d2[0] - d1[0] * d2[2],
                    d1[0] * d2[1] - d1[1] * d2[0]};
    g.area[f] = std::sqrt((cd[0] * cd[0] + cd[1] * cd[1]) + cd[2] * cd[2]) / 2.0;
    // FP64 frame [b t n].
    double nd[3] = {(double)n[0], (double)n[1], (double)n[2]};
    double nn = std::sqrt((nd[0] * nd[0] + nd[1] * nd[1]) + nd[2] * nd[2]);
    double nu[3] = {nd[0] / nn, nd[1] / nn, nd[2] / nn};
    double td[3] = {(double)(v[1][0] - v[0][0]), (double)(v[1][1] - v[0][1]),
                    (double)(v[1][2] - v[0][2])};
    double tn = std::sqrt((td[0] * td[0] + td[1] * td[1]) + td[2] * td[2]);
    double tu[3] = {td[0] / tn, td[1] / tn, td[2] / tn};
    double bu[3] = {tu[1] * nu[2] - tu[2] * nu[1], tu[2] * nu[0] - tu[0] * nu[2],
                    tu[0] * nu[1] - tu[1] * nu[0]};
    double *fr = &g.frame[9 * f];
    fr[0] = bu[0]; fr[1] = bu[1]; fr[2] = bu[2];
    fr[3] = tu[0]; fr[4] = tu[1]; fr[5] = tu[2];
    fr[6] = nu[0]; fr[7] = nu[1]; fr[8] = nu[2];
  }
  };
  const int64_t chunk = 1 << 16;
  const int nt = (int)std::max<int64_t>(
      1, std::min<int64_t>((nf + chunk - 1) / chunk, std::thread::hardware_concurrency()));
  if (nt == 1) {
    facets(0, nf);
  } else {
    std::vector<std::thread> pool;
    const int64_t per = (nf + nt - 1) / nt;
    for (int t = 0; t < nt; ++t)
      pool.emplace_back(facets, std::min(nf, t * per), std::min(nf, (t + 1) * per));
    for (auto &th : pool) th.join();
  }
  return g;
}

// Receiver clusters for the form-factor facing cull: facets [64c, 64c + 64)
// form cluster c, stored as (centre, threshold).  For a source j with FP32
// centroid p_j and normal n_j, when
//   t = ((cx - pjx) * njx + (cy - pjy) * njy) + (cz - pjz) * njz   (FP32)
// satisfies t <= -threshold, the FP32 facing test (p_i - p_j).n_j <= -r_i
// holds for every member i, so the whole cluster is culled without testing
// its members.  Exact bound: (p_i - p_j).n_j + r_i <= (c - p_j).n_j +
// |n_j| (|p_i - c| + r_i), with |n_j| <= 1 + 2^-20 for an FP32 unit normal.
// Rounding: each FP32 dot product of differences of coordinates bounded by E
// (the largest centroid coordinate) errs by at most 4u|d||n| <= 14uE
// (u = 2^-24, |d| <= 2 sqrt(3) E), so the pair test and t together err by
// less than 28uE; the threshold adds 64uE.  Facet order decides how tight
// the clusters are (index-local meshes such as subdivided icosahedra give
// compact clusters); a loose cluster is simply never culled.
std::vector<float> cull_clusters(const HostGeometry &g) {
  const int64_t nf = g.nf;
  const int64_t nc = (nf + 63) / 64;
  std::vector<float> out(4 * nc);
  double E = 0.0;
  for (int64_t k = 0; k < 3 * nf; ++k) E = std::max(E, std::fabs((double)g.cen[k]));
  const double margin = 64.0 * std::ldexp(1.0, -24) * E;
  for (int64_t c = 0; c < nc; ++c) {
    const int64_t i0 = 64 * c, i1 = std::min<int64_t>(nf, i0 + 64);
    double m[3] = {0, 0, 0};
    for (int64_t i = i0; i < i1; ++i)
      for (int a = 0; a < 3; ++a) m[a] += g.cen[3 * i + a];
    float cf[3];
    for (int a = 0; a < 3; ++a) cf[a] = (float)(m[a] / (double)(i1 - i0));
    double R = 0.0;
    for (int64_t i = i0; i < i1; ++i) {
      double d2 = 0.0;
      for (int a = 0; a < 3; ++a) {
        const double d = (double)g.cen[3 * i + a] - (double)cf[a];
        d2 += d * d;
      }
      R = std::max(R, std::sqrt(d2) + (double)g.rad[i]);
    }
    const double thr = R * (1.0 + std::ldexp(1.0, -20)) + margin;
    float tf = (float)thr;
    if ((double)tf < thr) tf = std::nextafter(tf, INFINITY);
    out[4 * c] = cf[0];
    out[4 * c + 1] = cf[1];
    out[4 * c + 2] = cf[2];
    out[4 * c + 3] = tf;
  }
  return out;
}

}  // namespace illum
