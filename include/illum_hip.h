/*
 * illum_hip.h -- C ABI of the MI355X (gfx950) form-factor / horizon / radiosity /
 * Crank-Nicolson library.
 *
 * Every entry point returns an int status (ILLUM_OK == 0).  On failure the
 * message is available from illum_last_error() (thread-local).  No entry point
 * exits the process or throws.  Buffers are caller-owned unless a handle says
 * otherwise; a pointer argument documented as "host or device" is read/written
 * on the device when the corresponding ILLUM_*_ON_DEVICE flag is set.
 *
 * Reference interfaces replaced (paths relative to the illum source tree):
 *   illum_mesh_*          obj_util::get_objects           src/cpp/obj_util.cpp:8-96
 *                         illum_context ctor               src/cpp/illum.cpp:22-27
 *   illum_bvh_build       Fast-BVH BVH{&objects} ctor      src/cpp/illum.cpp:25
 *   illum_form_factors    illum_context::compute_F         src/cpp/illum.cpp:31-237, illum.hpp:18-20
 *   illum_horizons        illum_context::make_horizons     src/cpp/illum.cpp:326-363, illum.hpp:22-23
 *   illum_direct_radiosity illum_context::get_direct_radiosity src/cpp/illum.cpp:382-466, illum.hpp:29-30
 *   illum_fib_spiral      fib_spiral                       src/cpp/illum.cpp:468-483, illum.hpp:50
 *   illum_radiosity_solve Gauss-Seidel loop of do_radiosity_task
 *                                                          src/cpp/illum_cli.cpp:425-440,
 *                                                          src/cpp/arma_util.cpp:3-19
 *   illum_thermal_*       thermal_model                    src/cpp/thermal.hpp:7-105
 *   illum_render          illum_renderer_cli's ray-cast    src/cpp/illum_renderer_cli.cpp:174-261
 *   illum_conductionQ     conductionQ (C++ wrapper)        src/cpp/conduction.hpp:16-28,
 *                                                          src/cpp/conduction.cpp:17-30
 */
#ifndef ILLUM_HIP_H
#define ILLUM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define ILLUM_OK                 0
#define ILLUM_ERR_ARG            1  /* bad argument (null, size, range)            */
#define ILLUM_ERR_NAN            2  /* a NaN form factor (illum.cpp:121-123)        */
#define ILLUM_ERR_ZERO_PIVOT     3  /* tridiagonal b(1) == 0 (tridag.for:10-12)     */
#define ILLUM_ERR_OOM            4  /* device or host allocation failed             */
#define ILLUM_ERR_HIP            5  /* HIP runtime error                            */
#define ILLUM_ERR_IO             6  /* file could not be read/parsed                */
#define ILLUM_ERR_CAPACITY       7  /* problem exceeds a documented size limit      */
#define ILLUM_ERR_NOT_CONVERGED  8  /* iterative solve hit max_iter above tolerance */
#define ILLUM_ERR_SHAPE          9  /* array of incompatible size (thermal.hpp:53-57) */

/* ---- flags ------------------------------------------------------------- */
#define ILLUM_IN_ON_DEVICE   1   /* input arrays are device pointers  */
#define ILLUM_OUT_ON_DEVICE  2   /* output arrays are device pointers */
#define ILLUM_ACCUMULATE     4   /* add into the output instead of overwriting it */
/* With every array on the device: return once the work is queued on the
 * context's stream instead of waiting for it.  Later calls on the same
 * context (illum_dev_copy included) see the results; illum_ctx_synchronize
 * waits for them.  A thermal zero pivot is then reported by the next
 * illum_thermal_get or synchronous step of that model.  Ignored where an
 * array is on the host. */
#define ILLUM_ASYNC          8
/* illum_form_factors_ex: also keep the raw visibility sets (the union of
 * closest hits per source before the zero drop) for
 * illum_sparse_copy_visibility_lists.  Testing aid; O(sets) device memory. */
#define ILLUM_FF_KEEP_VISIBILITY 16
/* illum_direct_radiosity: every body of a multi-body mesh has its own S suns,
 * sun_km is 3 x (S nbody) column-major with body b's suns in columns
 * b S .. b S + S - 1.  No reference counterpart (one body per process). */
#define ILLUM_SUN_PER_BODY   32
/* illum_timeloop_create: queue each step's kernels directly instead of
 * replaying captured HIP graphs (same results; for comparison). */
#define ILLUM_TL_EAGER       64
/* illum_thermal_run: Qseries is one row of nf values, the flux of every one
 * of the nsteps steps (the same steps as nsteps calls of illum_thermal_step
 * with that Q, run in one launch sequence).  No reference counterpart. */
#define ILLUM_Q_HELD         128

/* radiosity solve methods */
#define ILLUM_RAD_JACOBI_FP64   1   /* FP64 Jacobi sweeps B <- E + F B                  */
#define ILLUM_RAD_MIXED         2   /* FP16 corrections with a per-column power-of-two
                                       scale, FP64 residual refinement               */

/* thermal precision modes */
#define ILLUM_THERM_FP64        1   /* FP64 Crank-Nicolson, FP64 Thomas                  */
#define ILLUM_THERM_MIXED       2   /* FP32 Thomas sweeps + FP64 refinement / surface    */

typedef struct illum_ctx     illum_ctx;
typedef struct illum_mesh    illum_mesh;
typedef struct illum_bvh     illum_bvh;
typedef struct illum_sparse  illum_sparse;
typedef struct illum_thermal illum_thermal;
typedef struct illum_timeloop illum_timeloop;

const char *illum_last_error(void);
int illum_version(void);

/* ---- context: one HIP device + one stream; calls on a context are serialised.
 * Handles made on a context (meshes, trees, matrices, thermal models) are
 * destroyed before it.  The context keeps the device buffers of destroyed
 * sparse matrices for reuse by the next assembly (up to 8 GiB), and its
 * scratch workspaces, until illum_ctx_destroy. */
int illum_ctx_create(int device, illum_ctx **out);
int illum_ctx_destroy(illum_ctx *ctx);
int illum_ctx_synchronize(illum_ctx *ctx);
/* Per-stage kernel timing with hipEvents recorded on the context's stream.
 * When enabled, every launch of a named stage accumulates its duration. */
int illum_ctx_set_profiling(illum_ctx *ctx, int enable);
int illum_ctx_reset_profile(illum_ctx *ctx);
/* Fills up to cap entries; names are written as NUL-separated list into
 * names (names_cap bytes).  *n receives the number of stages. */
int illum_ctx_get_profile(illum_ctx *ctx, int cap, double *total_ms, int64_t *launches,
                          char *names, int names_cap, int *n);

/* The autotuner's record on this context: one line per tuned launch shape,
 * "key choice best_ms c1:ms1 c2:ms2 ...".  Launch shapes that do not change
 * results (the direct kernel's lanes per facet-sun pair, the wide SpMM's
 * right-hand sides per lane, the conduction buckets' stream layout) are
 * timed on the live inputs the first time a shape is seen and the fastest is
 * kept for the context's life; ILLUM_NO_AUTOTUNE=1 keeps the built-in
 * heuristics.  The BVH leaf size and walk are tuned by illum_bvh_build.
 * Writes at most cap bytes (NUL-terminated) and the full length to *len.
 * No reference counterpart. */
int illum_ctx_tuning(illum_ctx *ctx, char *out, int cap, int *len);

/* ---- device memory on the context's GPU (for the *_ON_DEVICE paths) -----
 * No counterpart in the reference (it has no device); lets C/C++ callers keep
 * E, B, horizons and Q series resident in HBM between calls.  kind: 0 host to
 * device, 1 device to host, 2 device to device; copies are ordered on the
 * context stream and complete before the call returns. */
int illum_dev_alloc(illum_ctx *ctx, int64_t bytes, void **out);
int illum_dev_free(illum_ctx *ctx, void *p);
int illum_dev_copy(illum_ctx *ctx, void *dst, const void *src, int64_t bytes, int kind);

/* ---- synthetic shape models (host, deterministic, seeded) --------------
 * Icosahedron subdivided `subdiv` times, projected to the unit sphere,
 * scaled to a triaxial ellipsoid (a, b, c), radially displaced by
 * `ncraters` Gaussian depressions with depth U(dmin,dmax), width U(wmin,wmax).
 * V: nv x 3 row-major doubles, F: nf x 3 row-major int32 (outward winding). */
int illum_synth_sizes(int subdiv, int64_t *nv, int64_t *nf);
int illum_synth_ellipsoid(int subdiv, double a, double b, double c, int ncraters,
                          double dmin, double dmax, double wmin, double wmax,
                          uint64_t seed, double *V, int32_t *F);

/* ---- meshes -------------------------------------------------------------
 * Vertices are rounded to FP32 and divided (in FP32) by vertex_divisor (100.0
 * reproduces the reference OBJ loader, obj_util.cpp:90-92; 1.0 = none); the
 * face normal is computed from the undivided FP32 vertices (obj_util.cpp:70).
 * When VN (vertex normals, indexed by vertex) is given, each face normal is
 * flipped to agree with the sum of its vertex normals (obj_util.cpp:64-88).
 * illum_mesh_from_obj loads the first shape of the file only (faces up to the
 * first "g"/"o" line that follows a face, obj_util.cpp:34), fan-triangulates
 * polygons and uses the "vn" normals when every face vertex has one. */
int illum_mesh_create(illum_ctx *ctx, const double *V, int64_t nv, const int32_t *F,
                      int64_t nf, const double *VN, double vertex_divisor,
                      illum_mesh **out);
int illum_mesh_from_obj(illum_ctx *ctx, const char *path, double vertex_divisor,
                        illum_mesh **out);
int illum_mesh_destroy(illum_mesh *mesh);
int illum_mesh_num_faces(const illum_mesh *mesh, int64_t *nf);
/* Partition the mesh into nbody independent bodies, facets
 * [body_face_off[b], body_face_off[b+1]) forming body b (offsets start at 0,
 * increase strictly and end at the facet count).  Visibility, form factors
 * and horizons of a facet then involve only its own body, so many bodies of
 * uneven size run in one launch; F comes out block-diagonal.  Set before
 * building BVHs (a tree built for another partition is rejected).  No
 * reference counterpart (the reference handles one body per process). */
int illum_mesh_set_bodies(illum_mesh *mesh, const int64_t *body_face_off, int nbody);
/* Derived per-facet geometry (host copies; any pointer may be null):
 * centroid (nf x 3 float), normal (nf x 3 float), bounding radius (nf float),
 * area (nf double), frame (nf x 9 double: columns b, t, n as in
 * get_frenet_frame, illum.cpp:244-262). */
int illum_mesh_geometry(const illum_mesh *mesh, float *centroid, float *normal,
                        float *radius, double *area, double *frame);
/* FP32 vertices of every facet after scaling, nf x 9 floats (v0, v1, v2). */
int illum_mesh_triangles(const illum_mesh *mesh, float *tri9);

/* ---- BVH (host SAH build; threaded, two-wide and four-wide layouts in HBM)
 * leaf_size 0 autotunes: the visibility kernel is timed (hipEvents) on a
 * strided sample of up to 8192 source facets for leaf sizes 2, 3, 4, 6, 8
 * and four walks -- per-lane near-first over the two-wide (layout 2) and
 * four-wide (4) trees, wavefront packets over the two-wide (3) and four-wide
 * (5) trees -- and the fastest tree and walk are kept; illum_bvh_tuning
 * reports the trials (leaf size, layout, ms).  A given leaf size uses the
 * two-wide packet walk.  Every walk gives the same visibility. */
int illum_bvh_build(illum_mesh *mesh, int leaf_size, illum_bvh **out);
int illum_bvh_tuning(const illum_bvh *bvh, int cap, int *leaf_sizes, int *layouts, double *ms,
                     int *n);
int illum_bvh_destroy(illum_bvh *bvh);
int illum_bvh_info(const illum_bvh *bvh, int64_t *nodes, int *depth, int *leaf_size);
/* The tree as the walks read it, copied to host arrays: sizes[4] = {threaded
 * nodes, two-wide inner nodes, triangles, bodies}, then (each array nullable;
 * a call with all of them null reports the sizes) nodes (8 floats per
 * threaded node: lo.xyz, miss link, hi.xyz, leaf info), wide (16 floats per
 * two-wide node: per child lo.xyz, reference, hi.xyz, 0), tris (12 floats per
 * triangle: v0.xyz, facet index bits, v1.xyz, 0, v2.xyz, 0) and roots (per
 * body: threaded, two-wide and four-wide root).  Measurement aid, no
 * reference counterpart: lets a CPU walk count the work of the same tree. */
int illum_bvh_export(const illum_bvh *bvh, int64_t *sizes, float *nodes, float *wide,
                     float *tris, int32_t *roots);
/* Closest-hit query for arbitrary rays (testing/diagnostics): origins and
 * directions are nray x 3 floats; hit receives the original facet index or
 * -1, t the parametric distance (in units of |d|). */
int illum_bvh_closest_hit(illum_bvh *bvh, const float *orig, const float *dir,
                          int64_t nray, int32_t *hit, float *t);

/* ---- form factors (A2 + A3) ---------------------------------------------
 * Builds the N x N sparse matrix F of compute_F: column j = source facet,
 * F(i,j) = mu_ij mu_ji A_j / (pi r_ij^2) over the union of closest hits of
 * rays from q_j = p_j + offset n_j towards every un-culled centroid,
 * exact zeros dropped, rows scaled by albedo (albedo_vec[i] if non-null,
 * else the scalar albedo).  Stored both as CSC (reference layout) and CSR. */
int illum_form_factors(illum_bvh *bvh, double albedo, const double *albedo_vec,
                       double offset, illum_sparse **out);
/* illum_form_factors with flags (0 or ILLUM_FF_KEEP_VISIBILITY).  Device
 * memory of an assembly is O(nnz): kept receivers are listed per source,
 * never as N x N bits.  A body may hold up to ~1.2M facets (the per-source
 * LDS bitset spans the source's body). */
int illum_form_factors_ex(illum_bvh *bvh, double albedo, const double *albedo_vec,
                          double offset, int flags, illum_sparse **out);
/* Work of one visibility pass over all sources, counted by a separate
 * instrumented launch (the assembly kernel carries no counters): out4 =
 * {rays, facing tests of candidate pairs after the cluster pass, ray/box
 * slab tests, ray/triangle tests}.  Measurement aid, no reference
 * counterpart. */
int illum_form_factors_work(illum_bvh *bvh, double offset, int64_t *out4);
/* The same launch with out[0..n), 1 <= n <= 8: the four counters above, then
 * the wavefront walks' own counts (zero for the per-lane walks): walks, node
 * visits and triangle tests per wavefront, and walks run with the triangle
 * test's axis permutation fixed for the whole wavefront.  Measurement aid, no
 * reference counterpart. */
int illum_form_factors_work_ex(illum_bvh *bvh, double offset, int64_t *out, int n);
int illum_sparse_create_csr(illum_ctx *ctx, int64_t n, int64_t nnz, const int64_t *rowptr,
                            const int32_t *colidx, const double *val, illum_sparse **out);
int illum_sparse_destroy(illum_sparse *sp);
int illum_sparse_info(const illum_sparse *sp, int64_t *n, int64_t *nnz);
int illum_sparse_copy_csc(const illum_sparse *sp, int64_t *colptr, int32_t *rowidx,
                          double *val);
int illum_sparse_copy_csr(const illum_sparse *sp, int64_t *rowptr, int32_t *colidx,
                          double *val);
/* Visibility sets of a matrix assembled with ILLUM_FF_KEEP_VISIBILITY: the
 * union of closest-hit facets per source before the zero drop
 * (illum.cpp:72-94).  As lists: hits_j = idx[ptr[j] .. ptr[j+1]) ascending
 * (ptr: nf + 1, idx: illum_sparse_visibility_nnz entries); or as nf rows of
 * ceil(nf/32) uint32 words, bit i of row j set <=> i in hits_j (small meshes).
 * Testing aids. */
int illum_sparse_visibility_nnz(const illum_sparse *sp, int64_t *nnz);
int illum_sparse_copy_visibility_lists(const illum_sparse *sp, int64_t *ptr, int32_t *idx);
int illum_sparse_visibility_words(const illum_sparse *sp, int64_t *words_per_row);
int illum_sparse_copy_visibility(const illum_sparse *sp, uint32_t *bits);
/* Number of visibility rays (facing-cull survivors) the assembly traced. */
int illum_sparse_rays(const illum_sparse *sp, int64_t *rays);

/* ---- horizons (A4) -------------------------------------------------------
 * out: nphi x nf column-major doubles (illum.hpp:47), host or device.
 * occluder may be null (use bvh); otherwise rays are cast against the
 * occluder mesh's BVH (illum.cpp:333-349). */
int illum_horizons(illum_bvh *bvh, illum_bvh *occluder, int nphi, double theta_eps,
                   double offset, double *out, int flags);

/* ---- direct radiosity (A5) -----------------------------------------------
 * horizons: nphi x nf column-major; sun_km: 3 x S column-major; disk_xy:
 * ndisk x 2 column-major (fib_spiral layout); E: nf x S column-major.
 * The reference picks a random disk tangent per facet (illum.cpp:406); this
 * library uses the deterministic tangent normalise(N x e_k), e_k the
 * coordinate axis least aligned with N. */
int illum_fib_spiral(double *xy, int n);
int illum_direct_radiosity(illum_mesh *mesh, const double *horizons, int nphi,
                           const double *sun_km, int S, const double *disk_xy,
                           int ndisk, double *E, int flags);

/* ---- renderer ray-cast (illum_renderer_cli.cpp:174-261) ----------------
 * Per-pixel closest-hit facet of the renderer's camera, height x width
 * row-major (row 0 = top image row): mode 0 orthographic at azimuth az_deg,
 * elevation el_deg; mode 1 equal-area cylindrical (az, el used as given, as
 * the reference does).  facet[p] = hit facet or -1; value[p] = data[facet]
 * clamped to [minval, maxval] (pass -inf / +inf for no clamp), NaN where
 * nothing is hit.  Either output may be null; data (nf doubles) is needed
 * for value.  Host or device per flags.  Colour mapping and image files are
 * left to the caller. */
int illum_render(illum_bvh *bvh, int mode, int width, int height, double az_deg,
                 double el_deg, const double *data, double minval, double maxval,
                 int32_t *facet, double *value, int flags);

/* ---- radiosity solve (A6) ------------------------------------------------
 * Solves (I - F) B = E for S right-hand sides; E, B are nf x S column-major
 * (host or device per flags).  Stops when max_s ||E - (I-F)B||_2/||E||_2
 * <= rel_tol.  resid (nullable, max_iter+1 doubles) receives the relative
 * residual after each FP64 check; *iters the number of F-products. */
int illum_radiosity_solve(illum_sparse *F, const double *E, int S, int method,
                          double rel_tol, int max_iter, double *B, double *resid,
                          int *iters, int flags);
/* illum_radiosity_solve that allocates no device memory.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  E and B
 * are device arrays (flags must hold ILLUM_IN_ON_DEVICE and
 * ILLUM_OUT_ON_DEVICE) and max_iter <= 256.  Every scratch buffer of the solve
 * (control block, residual history, iterates, FP16 corrections, partial sums)
 * comes from one fixed 64 MiB array per GPU instead of the context's
 * workspace; a solve that needs more (about nf x S > 1.3 million values for
 * the mixed method, 2.4 million for FP64) returns ILLUM_ERR_CAPACITY before
 * queuing anything.  All calls on one GPU share that array and it keeps the
 * last call's contents (every solve initialises what it reads): calls that
 * overlap on one GPU, from different contexts or threads, race.  Results are
 * identical to illum_radiosity_solve's. */
int illum_radiosity_solve_noalloc(illum_sparse *F, const double *E, int S, int method,
                                  double rel_tol, int max_iter, double *B, double *resid,
                                  int *iters, int flags);

/* ---- thermal model (A7 + A8 + A9) --------------------------------------
 * ngrid layer grids: grid g has depths z[grid_off[g] .. grid_off[g+1]) (at
 * least 2 layers, strictly increasing, z > 0) with per-layer thermal inertia
 * ti and volumetric heat capacity rhoc at the same offsets.  Facet f uses
 * grid grid_of_facet[f] (null => all use grid 0) and scales its thermal
 * inertia by ti_scale[f] (null => 1).  T0: scalar initial temperature if
 * T0_full is null, else T0_full is (nzmax+1) x nf column-major with row 0 the
 * surface (thermal.hpp:47-58).  emissivity: thermal.hpp:77 uses 0.999. */
int illum_thermal_create(illum_ctx *ctx, int64_t nf, int ngrid, const int64_t *grid_off,
                         const double *z, const double *ti, const double *rhoc,
                         const int32_t *grid_of_facet, const double *ti_scale,
                         double emissivity, double T0, const double *T0_full,
                         int precision, illum_thermal **out);
int illum_thermal_destroy(illum_thermal *th);
int illum_thermal_info(const illum_thermal *th, int64_t *nf, int *nzmax);
/* One step of thermal_model::step(dt, Q) (thermal.hpp:65-105): Q is nf
 * doubles (host or device), Qprev is the previous step's Q (zero initially). */
int illum_thermal_step(illum_thermal *th, double dt, const double *Q, int flags);
/* nsteps consecutive steps, Q series nsteps x nf (row k = Q at step k+1),
 * run as one persistent launch sequence (ILLUM_Q_HELD: one row for all). */
int illum_thermal_run(illum_thermal *th, double dt, const double *Qseries, int nsteps,
                      int flags);
/* T: (nzmax+1) x nf column-major (ragged facets padded with NaN), Tsurf and
 * Fsurf: nf; host or device per flags; any may be null. */
int illum_thermal_get(illum_thermal *th, double *T, double *Tsurf, double *Fsurf,
                      int flags);
/* thermal_model::get_radiosity (thermal.hpp:61-63): out = -Fsurf (nf
 * doubles), or out += -Fsurf with ILLUM_ACCUMULATE (the rad = direct -
 * Fsurf sum of illum_cli.cpp:411-423 without leaving the device). */
int illum_thermal_get_radiosity(illum_thermal *th, double *out, int flags);

/* ---- coupled radiosity + thermal time loop (illum_cli.cpp:403-456) ------
 * One step: E = direct insolation under the step's sun minus the thermal
 * model's surface flux (rad = -F_surf + direct, illum_cli.cpp:411-423), the
 * solve (I - F) B = E by `method` to rel_tol, then one Crank-Nicolson step
 * of th with Q = B (illum_cli.cpp:446).  Everything stays on the device and a
 * run waits on the host only at its end: each solve queues a fixed budget of
 * `products` products (0: two more than a checked solve of the first step
 * takes, at least 4), products after convergence return at once, and a
 * solve still above rel_tol after its budget is counted.  Steps are captured
 * once into HIP graphs (1 and 8 steps) and replayed; ILLUM_TL_EAGER queues
 * the same kernels directly.  horizons (nphi x nf column-major) and disk_xy
 * (ndisk x 2 column-major) are device arrays that must outlive the loop; F,
 * mesh and th belong to one context and must outlive it too.  The
 * reference's fixed Gauss-Seidel sweep count (gs_steps) is replaced by the
 * residual stop. */
int illum_timeloop_create(illum_sparse *F, illum_mesh *mesh, illum_thermal *th,
                          const double *horizons, int nphi, const double *disk_xy, int ndisk,
                          int method, double rel_tol, int products, int flags,
                          illum_timeloop **out);
/* nsteps steps with suns sun_km (3 x nsteps column-major, km; host, or
 * device with ILLUM_IN_ON_DEVICE) and time step dt.  Returns
 * ILLUM_ERR_NOT_CONVERGED when a solve ended above rel_tol (the steps still
 * ran) and ILLUM_ERR_ZERO_PIVOT after a zero pivot.  With ILLUM_IN_ON_DEVICE
 * | ILLUM_ASYNC it returns once the steps are queued (illum_timeloop_stats
 * then waits and reports). */
int illum_timeloop_run(illum_timeloop *tl, const double *sun_km, int nsteps, double dt,
                       int flags);
/* steps run so far; the product budget; over the last run: the most products
 * a solve took, the solves that ended above rel_tol, the largest final
 * relative residual.  Any pointer may be null. */
int illum_timeloop_stats(illum_timeloop *tl, int64_t *steps, int *products, int *max_products,
                         int *unconverged, double *max_resid);
/* E and B of the last step (nf doubles each; host, or device with
 * ILLUM_OUT_ON_DEVICE); either may be null. */
int illum_timeloop_get(illum_timeloop *tl, double *E, double *B, int flags);
int illum_timeloop_destroy(illum_timeloop *tl);

/* ---- mesh-field transfer (src/py/interpolate_mesh_field.py,
 * src/py/build_mesh_transfer_tables.py) ----------------------------------
 * illum_knn: for each of nt target points the k (<= 32) nearest of ns source
 * points (both n x 3 row-major FP64), nearest first: idx nt x k, dist
 * (nullable) nt x k Euclidean distances; the reference queries a scipy
 * cKDTree.  Exact brute force in FP64, ties to the smaller source index.
 * Host or device arrays per flags. */
int illum_knn(illum_ctx *ctx, const double *src, int64_t ns, const double *tgt, int64_t nt,
              int k, int32_t *idx, double *dist, int flags);
/* The interpolation of interpolate_mesh_field.py on host arrays: Xsrc is
 * nlayers x ns row-major, Xtgt nlayers x nt.  mode 0 nearest, 1 weighted
 * (the reference's weights D / sum(D) over the k nearest), 2 area_weighted
 * (the same weights on X times the source areas, divided by the target's
 * area; area_src and area_tgt required). */
int illum_mesh_transfer(illum_ctx *ctx, const double *src, int64_t ns, const double *tgt,
                        int64_t nt, const double *Xsrc, int nlayers, int mode, int k,
                        const double *area_src, const double *area_tgt, double *Xtgt);

/* Single-facet drop-in for conductionQ (conduction.hpp:16-28), run on the
 * GPU.  T (nz) and Tsurf, Fsurf are updated in place. */
int illum_conductionQ(int nz, const double *z, double dt, double Qn, double Qnp1,
                      double *T, const double *ti, const double *rhoc, double emiss,
                      double *Tsurf, double Fgeotherm, double *Fsurf);

#ifdef __cplusplus
}
#endif

#endif /* ILLUM_HIP_H */
