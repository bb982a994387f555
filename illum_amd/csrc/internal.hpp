// Internal types shared by the host-side C++ and the HIP translation units.
#pragma once

#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>
#include <map>
#include <memory>
#include <utility>
#include <stdexcept>

#include "../../include/illum_hip.h"

namespace illum {

// Status-carrying exception used inside the library; every C entry point
// catches it and converts it to (status, illum_last_error()).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

// Allocator whose value-initialisation leaves elements uninitialised, so
// vector::resize does not zero (and fault in) hundreds of MB on one thread
// for arrays that are then written in full by the threads computing them.
template <class T>
struct NoInitAlloc : std::allocator<T> {
  template <class U> struct rebind { using other = NoInitAlloc<U>; };
  NoInitAlloc() = default;
  template <class U> NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
  template <class U> void construct(U *p) noexcept { ::new ((void *)p) U; }
  template <class U, class... A> void construct(U *p, A &&...a) {
    ::new ((void *)p) U(std::forward<A>(a)...);
  }
};
template <class T> using HostVec = std::vector<T, NoInitAlloc<T>>;

// ---- host-side geometry (built once per mesh, uploaded) ------------------
struct HostGeometry {
  int64_t nf = 0;
  HostVec<float> tri;      // nf x 9 : v0, v1, v2 (FP32, after divisor)
  HostVec<float> cen;      // nf x 3 : centroid
  HostVec<float> nrm;      // nf x 3 : unit normal (FP32)
  HostVec<float> rad;      // nf     : bounding radius
  HostVec<double> area;    // nf     : |(v1-v0)x(v2-v0)|/2 in FP64
  HostVec<double> frame;   // nf x 9 : b, t, n columns (FP64)
  // Bodies: facets [body_off[b], body_off[b+1]) form body b, a scene of its
  // own (empty or {0, nf}: one body).
  std::vector<int64_t> body_off;
};

// Builds HostGeometry from double vertices (rounded to float) and faces.
HostGeometry make_geometry(const double *V, int64_t nv, const int32_t *F, int64_t nf,
                           const double *VN, double vertex_divisor);

// Receiver clusters of 64 consecutive facets for the form-factor facing cull:
// (centre, threshold) float4 per cluster; the bound is documented in mesh.cpp.
std::vector<float> cull_clusters(const HostGeometry &g);

// Synthetic cratered ellipsoid.
void synth_sizes(int subdiv, int64_t *nv, int64_t *nf);
void synth_ellipsoid(int subdiv, double a, double b, double c, int ncraters, double dmin,
                     double dmax, double wmin, double wmax, uint64_t seed, double *V,
                     int32_t *F);

// Minimal Wavefront OBJ reader (v / vn / f, polygon faces fan-triangulated).
void read_obj(const char *path, std::vector<double> &V, std::vector<int32_t> &F,
              std::vector<double> &VN);

// ---- flattened threaded BVH ----------------------------------------------
// Node i occupies two float4 slots:
//   lo = (bmin.x, bmin.y, bmin.z, as_float(miss))  miss = next node when the
//                                                   box is missed or the leaf
//                                                   is done (-1 = end)
//   hi = (bmax.x, bmax.y, bmax.z, as_float(info))  info = -1 for an inner node
//                                                   (hit link = i + 1), else
//                                                   (count << 24) | first
// Triangles are stored in leaf order as three float4 each:
//   (v0.xyz, as_float(orig_index)), (v1.xyz, 0), (v2.xyz, 0).
//
// The same tree is also stored in a two-wide layout for near-first traversal:
// inner node k (pre-order over inner nodes only) holds both children's boxes,
//   f4[0] = (left.bmin, as_float(left)),  f4[1] = (left.bmax, 0),
//   f4[2] = (right.bmin, as_float(right)), f4[3] = (right.bmax, 0),
// where a child reference >= 0 is an inner node index and < 0 is ~leaf_info.
// A tree whose root is a leaf gets one inner root with an empty right box.
struct HostBVH {
  std::vector<float> nodes;    // 8 floats per node
  std::vector<float> wide;     // 16 floats per inner node
  std::vector<float> wide4;    // 32 floats per four-wide node
  std::vector<float> tris;     // 12 floats per triangle
  int64_t nnodes = 0;
  int64_t ninner = 0;
  int64_t ninner4 = 0;
  int depth = 0;
  int depth4 = 0;
  int leaf_size = 0;
  std::vector<int32_t> root_thr, root_wide, root_wide4;  // per body (one for one body)
};
HostBVH build_bvh(const HostGeometry &g, int leaf_size);

}  // namespace illum

// ---- opaque handle definitions ------------------------------------------
struct illum_ctx {
  int device = 0;
  void *stream = nullptr;              // hipStream_t
  // Device address of the no-allocation solve's fixed scratch on this
  // context's GPU (resolved at context creation).
  void *noalloc_scratch = nullptr;
  // Auxiliary streams (default priority) and their join events, made on
  // first use, for independent launches forked from `stream` and joined
  // back into it within one call (the conduction depth buckets).
  std::vector<void *> aux_streams, aux_events;
  void *fork_event = nullptr;
  bool profiling = false;
  struct Stage { double ms = 0; int64_t launches = 0; };
  std::map<std::string, Stage> profile;
  std::vector<std::pair<std::string, std::pair<void *, void *>>> pending;  // events
  // Grow-only device scratch kept between calls (slot -> pointer, bytes), so
  // repeated calls skip large allocations.  Calls on one context are issued
  // from one host thread at a time, in stream order.
  std::vector<std::pair<void *, size_t>> workspace;
  // Reuse pool for the buffers of form-factor matrices: a destroyed
  // matrix's blocks stay here and the next assembly takes them (best fit),
  // so a loop of assemblies allocates and frees nothing.  Blocks are reused
  // in stream order on this context's stream.
  // Autotuner record: choices of launch shapes that do not change results,
  // timed on the live inputs the first time a shape is seen (key -> choice),
  // and every trial for reporting.
  struct Tuned { std::string key; int choice; double best_ms; std::string trials; };
  std::vector<Tuned> tuned;
  std::map<std::string, int> tune_cache;
  std::multimap<size_t, void *> pool;   // capacity -> free block
  std::map<void *, size_t> pool_cap;    // every block from pool_alloc -> capacity
  size_t pool_bytes = 0;                // bytes held free in `pool`
};

struct illum_mesh {
  illum_ctx *ctx = nullptr;
  illum::HostGeometry g;
  // device copies
  float *d_pr4 = nullptr;      // nf x float4 (centroid, radius)
  float *d_n4 = nullptr;       // nf x float4 (normal, 0)
  double *d_area = nullptr;    // nf
  double *d_frame = nullptr;   // nf x 9
  float *d_clus = nullptr;     // ceil(nf / 64) x float4 (cull_clusters)
  int64_t nclus = 0;
  // bodies (see HostGeometry::body_off): facet ranges and, with several
  // bodies, the body of each facet
  int nbody = 1;
  int32_t *d_boff = nullptr;   // nbody + 1
  int32_t *d_body = nullptr;   // nf, null for one body
  // Direct insolation's queue of open (facet, sun) pairs for its two-pass
  // shape: [0] count, [1] finished workgroups of the second pass, then
  // dq_cap pair indices.  Zeroed at creation; the second pass leaves it
  // zeroed again.
  int32_t *d_dq = nullptr;
  int64_t dq_cap = 0;
};

struct illum_bvh {
  illum_mesh *mesh = nullptr;
  illum::HostBVH h;
  float *d_nodes = nullptr;
  float *d_tris = nullptr;
  float *d_wide = nullptr;
  float *d_wide4 = nullptr;
  int32_t *d_roots = nullptr;  // per body: threaded, two-wide, four-wide root, 0
  int nbody = 1;               // bodies of the mesh when the tree was built
  std::vector<int64_t> body_off;  // the partition the forest was built for
  // form-factor traversal (2 / 4 = two- / four-wide near-first walk per lane,
  // 3 / 5 = wavefront packet walk over the two- / four-wide tree); without
  // autotuning the two-wide packet walk
  int ff_layout = 3;
  // autotuning record (illum_bvh_build with leaf_size 0): leaf size x layout
  int ntuned = 0;
  int tune_leaf[32] = {0};
  int tune_layout[32] = {0};
  double tune_ms[32] = {0};
  // nonzeros of the last form-factor assembly on this tree (sizes the next
  // one's pool), 0 before the first
  int64_t ff_nnz_last = 0;
  // set once more than 1/64 of the sources of a paged visibility pass on
  // this tree ran out of pool pages: later passes keep a bitset
  bool ff_paged_off = false;
};

constexpr int64_t CSR_PAD = 64;

struct illum_sparse {
  illum_ctx *ctx = nullptr;
  int64_t n = 0, nnz = 0;
  bool has_csc = false;
  int64_t *d_colptr = nullptr;
  int32_t *d_rowidx = nullptr;
  double *d_cval = nullptr;
  // CSR; colidx and rval hold CSR_PAD entries past nnz (zero), so a kernel
  // may read a whole group of nonzeros past a row's end
  int64_t *d_rowptr = nullptr;
  int32_t *d_colidx = nullptr;
  double *d_rval = nullptr;
  // raw visibility sets (union of first hits, before the zero drop), kept
  // on request: source j's hits are d_vidx[d_vptr[j] .. d_vptr[j+1]), ascending
  int64_t vnnz = 0;
  int64_t *d_vptr = nullptr;
  int32_t *d_vidx = nullptr;
  // visibility-kernel work: rays, pair tests, box tests, triangle tests
  unsigned long long *d_nrays = nullptr;
  // products the previous solve took: the next solve's first batch
  int last_nprod = 0;
};

struct illum_thermal {
  illum_ctx *ctx = nullptr;
  int64_t nf = 0;
  // Facets are permuted into "slots": each 64-slot wavefront holds facets of
  // one layer grid (padding slots hold -1); per-layer state is stored
  // slot-major, T[l * nslot + slot].
  int64_t nslot = 0;
  int32_t *d_fidx = nullptr;   // slot -> facet (or -1)
  int32_t *d_pos = nullptr;    // facet -> slot
  int32_t *d_wgrid = nullptr;  // wavefront -> grid
  struct Launch { int nz_bucket; int64_t block0, nblocks; };
  std::vector<Launch> launches;  // per layer-count bucket (0 = streaming kernel)
  int nzmax = 0;
  int precision = ILLUM_THERM_FP64;
  int ngrid = 0;
  double emiss = 0.999;
  int32_t *d_nz = nullptr;     // per facet layer count
  int32_t *d_grid = nullptr;   // per facet grid id
  double *d_s2 = nullptr;      // per facet ti_scale^2
  void *impl = nullptr;        // host tables + per-dt device tables (thermal.hip)
  int64_t *d_goff = nullptr;
  double *d_T = nullptr;       // (nzmax+1) x nslot, row 0 = surface
  double *d_Qprev = nullptr;   // nf
  double *d_F = nullptr;       // nf surface flux
  double *d_work = nullptr;    // streaming-kernel scratch ((nzmax+1) x nslot), if needed
  float *d_work32 = nullptr;   // streaming-kernel scratch (2 (nzmax+1) x nslot), if needed
  double *d_Qdev = nullptr;    // staging for host Q
  int64_t qdev_cap = 0;
  int *d_status = nullptr;
  // grid tables on host (for validation / get)
  std::vector<int64_t> goff;
  std::vector<int32_t> nz_host;
};

namespace illum {
// A forest is only valid for the body partition it was built for.
inline void check_bvh_bodies(const illum_bvh *bvh) {
  if (bvh->body_off != bvh->mesh->g.body_off)
    throw Error(ILLUM_ERR_ARG, "the BVH was built for another body partition; rebuild it");
}
}  // namespace illum
