"""TEST INFRASTRUCTURE ONLY: Python binding of the CPU oracle (liboracle.so)
and of the reference's own Fortran conductionQ compiled into _ref/.

Imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, always as the checker.  See oracle.cpp for the reference file:line each
function restates.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_conduction.so")

_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
_D = C.c_double

_SIG = {
    "oracle_geometry": ([_P, _I64, _P, _I64, _P, _D, _P, _P, _P, _P, _P, _P], _I),
    "oracle_closest_hit": ([_P, _I64, _P, _P, _I64, _I, _I, _P, _P], _I),
    "oracle_visibility": ([_P, _P, _P, _P, _I64, _D, _I64, _I64, _I, _I, _P], _I64),
    "oracle_walk_work": ([_P, _P, _P, _I64, _D, _I64, _I64, _P, _I, _P, _I, _P, _I, C.c_float,
                          _I, _P, _P], _I),
    "oracle_form_factors": ([_P, _P, _P, _I64, _P, _I64, _I64, _D, _P, _I64, _P, _P, _P], _I64),
    "oracle_horizons": ([_P, _P, _P, _I64, _I64, _P, _I64, _I, _D, _D, _I, _I, _P], _I64),
    "oracle_fib_spiral": ([_P, _I], _I),
    "oracle_direct": ([_P, _P, _I64, _P, _I, _P, _I, _P, _I, _I, _P], _I),
    "oracle_gauss_seidel": ([_I64, _P, _P, _P, _P, _I, _D, _P, _P], _I),
    "oracle_conductionQ": ([_I, _P, _D, _D, _D, _P, _P, _P, _D, _P, _D, _P], _I),
    "oracle_thermal_run": ([_I64, _P, _P, _P, _P, _P, _P, _D, _I, _D, _I, _P, _P, _P, _P, _P, _I],
                           _I),
    "oracle_render": ([_P, _P, _I64, _I, _I, _I, _D, _D, _I, _P], _I),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} not built (make -C oracle)")
        L = C.CDLL(LIB)
        for k, (a, r) in _SIG.items():
            getattr(L, k).argtypes = a
            getattr(L, k).restype = r
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def nthreads_default():
    return max(1, min(16, os.cpu_count() or 1))


def geometry(V, F, VN=None, divisor=1.0):
    V = np.ascontiguousarray(V, np.float64)
    F = np.ascontiguousarray(F, np.int32)
    nf = F.shape[0]
    out = dict(tri=np.zeros((nf, 9), np.float32), centroid=np.zeros((nf, 3), np.float32),
               normal=np.zeros((nf, 3), np.float32), radius=np.zeros(nf, np.float32),
               area=np.zeros(nf), frame=np.zeros((nf, 9)))
    VN = None if VN is None else np.ascontiguousarray(VN, np.float64)
    st = lib().oracle_geometry(_p(V), V.shape[0], _p(F), nf, _p(VN), divisor, _p(out["tri"]),
                               _p(out["centroid"]), _p(out["normal"]), _p(out["radius"]),
                               _p(out["area"]), _p(out["frame"]))
    if st:
        raise ValueError("oracle_geometry failed")
    return out


def read_obj_first_shape(path):
    """The OBJ input of obj_util::get_objects (obj_util.cpp:8-96) with
    shape_index 0: returns (V, F, VN) where F holds the triangles of the
    first shape only (a shape ends at a "g" or "o" line that follows faces,
    as tinyobjloader splits them), polygons fan-triangulated, face vertices
    taken before any "/" (normal indices in faces are not used: the reference
    reads normals[3 * vertex_index], obj_util.cpp:71-85).  V and VN are the
    whole file's pools; VN is None unless it has a normal for every vertex
    the first shape uses."""
    V, VN, shapes, cur = [], [], [], []
    for line in open(path):
        tok = line.split()
        if not tok or tok[0].startswith("#"):
            continue
        if tok[0] in ("g", "o"):
            if cur:
                shapes.append(cur)
                cur = []
        elif tok[0] == "v":
            V.append([float(x) for x in tok[1:4]])
        elif tok[0] == "vn":
            VN.append([float(x) for x in tok[1:4]])
        elif tok[0] == "f":
            idx = []
            for t in tok[1:]:
                k = int(t.split("/")[0])
                idx.append(k - 1 if k > 0 else len(V) + k)
            cur += [[idx[0], idx[t], idx[t + 1]] for t in range(1, len(idx) - 1)]
    if cur:
        shapes.append(cur)
    V = np.asarray(V, np.float64)
    F = np.asarray(shapes[0], np.int32)
    VN = np.asarray(VN, np.float64) if len(VN) > F.max() else None
    return V, F, VN


def get_objects(path, divisor=100.0):
    """Geometry of obj_util::get_objects + Tri (obj_util.cpp:43-93): the first
    shape, face normals oriented by the vertex normals, vertices / 100."""
    V, F, VN = read_obj_first_shape(path)
    return geometry(V, F, VN=VN, divisor=divisor)


def closest_hit(tri, orig, dirs, use_bvh=True, nthreads=None):
    tri = np.ascontiguousarray(tri, np.float32)
    orig = np.ascontiguousarray(orig, np.float32)
    dirs = np.ascontiguousarray(dirs, np.float32)
    n = orig.shape[0]
    hit = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    lib().oracle_closest_hit(_p(tri), tri.shape[0], _p(orig), _p(dirs), n, int(use_bvh),
                             nthreads or nthreads_default(), _p(hit), _p(t))
    return hit, t


def render(g, mode=0, width=64, height=64, az=0.0, el=0.0, nthreads=None):
    """Per-pixel hit facet of the renderer's ray-cast (height x width, -1 = none)."""
    out = np.zeros((height, width), np.int32)
    lib().oracle_render(_p(g["tri"]), _p(g["centroid"]), g["tri"].shape[0], mode, width, height,
                        az, el, nthreads or nthreads_default(), _p(out))
    return out


def visibility(g, offset=1e-5, j0=0, j1=None, use_bvh=True, nthreads=None):
    """Raw visibility rows (union of closest hits) for sources [j0, j1)."""
    nf = g["tri"].shape[0]
    j1 = nf if j1 is None else j1
    words = (nf + 31) // 32
    bits = np.zeros((j1 - j0, words), np.uint32)
    nr = lib().oracle_visibility(_p(g["tri"]), _p(g["centroid"]), _p(g["normal"]),
                                 _p(g["radius"]), nf, offset, j0, j1, int(use_bvh),
                                 nthreads or nthreads_default(), _p(bits))
    return bits, nr


def walk_work(g, tree, body=0, offset=1e-5, j0=0, j1=None, cap=8, tinit=1.0009765625,
              nthreads=None, with_bits=False):
    """Box and triangle tests of the form-factor rays of sources [j0, j1)
    walked one ray at a time over the product's exported tree (the dict of
    BVH.export()): {"rays", "box_tests", "tri_tests"} and, with with_bits,
    the visibility rows.  cap: the per-lane walk's stack entries (deeper
    walks are answered by the threaded walk); tinit: its first bound."""
    nf = g["tri"].shape[0]
    j1 = nf if j1 is None else j1
    words = (nf + 31) // 32
    bits = np.zeros((j1 - j0, words), np.uint32) if with_bits else None
    out = np.zeros(3, np.int64)
    nodes = np.ascontiguousarray(tree["nodes"], np.float32)
    wide = np.ascontiguousarray(tree["wide"], np.float32)
    tris = np.ascontiguousarray(tree["tris"], np.float32)
    roots = tree["roots"]
    lib().oracle_walk_work(_p(g["centroid"]), _p(g["normal"]), _p(g["radius"]), nf, offset,
                           j0, j1, _p(nodes), int(roots[body, 0]), _p(wide),
                           int(roots[body, 1]), _p(tris), cap, tinit,
                           nthreads or nthreads_default(), _p(out), _p(bits))
    res = {"rays": int(out[0]), "box_tests": int(out[1]), "tri_tests": int(out[2])}
    return (res, bits) if with_bits else res


def form_factors(g, bits, j0=0, j1=None, albedo=1.0, albedo_vec=None):
    """CSC (colptr, rowidx, val) of columns [j0, j1) from visibility rows."""
    nf = g["tri"].shape[0]
    j1 = nf if j1 is None else j1
    cap = int(bits.astype(np.uint64).sum() and np.unpackbits(bits.view(np.uint8)).sum()) + 1
    colptr = np.zeros(j1 - j0 + 1, np.int64)
    rowidx = np.zeros(cap, np.int32)
    val = np.zeros(cap)
    av = None if albedo_vec is None else np.ascontiguousarray(albedo_vec, np.float64)
    nnz = lib().oracle_form_factors(_p(g["centroid"]), _p(g["normal"]), _p(g["area"]), nf,
                                    _p(np.ascontiguousarray(bits)), j0, j1, albedo, _p(av), cap,
                                    _p(colptr), _p(rowidx), _p(val))
    if nnz == -2:
        raise FloatingPointError("computed a NaN form factor")
    if nnz < 0:
        raise RuntimeError("oracle_form_factors capacity")
    return colptr, rowidx[:nnz].copy(), val[:nnz].copy()


def horizons(g, nphi=361, theta_eps=1e-3, offset=1e-5, j0=0, j1=None, occ_tri=None,
             use_bvh=True, nthreads=None):
    """nphi x (j1 - j0) horizon angles."""
    nf = g["tri"].shape[0]
    j1 = nf if j1 is None else j1
    occ = g["tri"] if occ_tri is None else np.ascontiguousarray(occ_tri, np.float32)
    out = np.zeros((j1 - j0, nphi))
    nr = lib().oracle_horizons(_p(g["centroid"]), _p(g["normal"]), _p(g["frame"]), j0, j1,
                               _p(occ), occ.shape[0], nphi, theta_eps, offset, int(use_bvh),
                               nthreads or nthreads_default(), _p(out))
    return out.T, nr


def fib_spiral(n):
    xy = np.zeros((2, n))
    lib().oracle_fib_spiral(_p(xy), n)
    return xy.T.copy()


def direct(g, hz, sun, disk_xy, nthreads=None):
    """E (N,) or (N, S) for sun (3,) or (3, S) km; hz nphi x N."""
    sun = np.asarray(sun, np.float64)
    single = sun.ndim == 1
    sunr = np.ascontiguousarray(sun.reshape(3, -1).T)
    S = sunr.shape[0]
    nf = g["tri"].shape[0]
    hzc = np.ascontiguousarray(np.asarray(hz).T)  # N x nphi row-major == nphi x N col-major
    disk = np.ascontiguousarray(np.asarray(disk_xy, np.float64).T)  # 2 x n == n x 2 col-major
    E = np.zeros((S, nf))
    lib().oracle_direct(_p(g["centroid"]), _p(g["frame"]), nf, _p(hzc), hzc.shape[1], _p(sunr),
                        S, _p(disk), disk.shape[1], nthreads or nthreads_default(), _p(E))
    return E[0] if single else E.T


def gauss_seidel(n, colptr, rowidx, val, E, max_sweeps=50, rel_tol=1e-10):
    """Reference GS; returns (B, sweeps, absolute residual per sweep)."""
    colptr = np.ascontiguousarray(colptr, np.int64)
    rowidx = np.ascontiguousarray(rowidx, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    E = np.ascontiguousarray(E, np.float64)
    B = np.zeros(n)
    res = np.zeros(max(1, max_sweeps))
    k = lib().oracle_gauss_seidel(n, _p(colptr), _p(rowidx), _p(val), _p(E), max_sweeps, rel_tol,
                                  _p(B), _p(res))
    return B, k, res[:k]


def conductionQ(nz, z, dt, Qn, Qnp1, T, ti, rhoc, emiss, Tsurf, Fgeotherm):
    T = np.array(T, np.float64)
    ts = C.c_double(Tsurf)
    fs = C.c_double(0.0)
    z = np.ascontiguousarray(z, np.float64)
    ti = np.ascontiguousarray(np.broadcast_to(ti, (nz,)), np.float64)
    rhoc = np.ascontiguousarray(np.broadcast_to(rhoc, (nz,)), np.float64)
    st = lib().oracle_conductionQ(nz, _p(z), dt, Qn, Qnp1, _p(T), _p(ti), _p(rhoc), emiss,
                                  C.byref(ts), Fgeotherm, C.byref(fs))
    if st:
        raise ZeroDivisionError("tridag: rewrite equations")
    return T, ts.value, fs.value


def thermal_run(goff, z, ti, rhoc, T, Qprev, Q, dt, emiss, grid=None, ti_scale=None,
                record=False, nthreads=None):
    """Runs Q.shape[0] steps; T is (nzmax+1) x N and is updated (a copy is returned)."""
    goff = np.ascontiguousarray(goff, np.int64)
    z = np.ascontiguousarray(z, np.float64)
    ti = np.ascontiguousarray(ti, np.float64)
    rhoc = np.ascontiguousarray(rhoc, np.float64)
    T = np.asfortranarray(np.array(T, np.float64))
    Qprev = np.array(Qprev, np.float64)
    Q = np.ascontiguousarray(Q, np.float64)
    nsteps, nf = Q.shape
    nzmax = T.shape[0] - 1
    Fs = np.zeros(nf)
    series = np.zeros((nsteps, nf)) if record else None
    g = None if grid is None else np.ascontiguousarray(grid, np.int32)
    s = None if ti_scale is None else np.ascontiguousarray(ti_scale, np.float64)
    st = lib().oracle_thermal_run(nf, _p(goff), _p(z), _p(ti), _p(rhoc), _p(g), _p(s), emiss,
                                  nzmax, dt, nsteps, _p(Q), _p(Qprev), T.ctypes.data_as(C.c_void_p),
                                  _p(Fs), _p(series), nthreads or nthreads_default())
    if st:
        raise ZeroDivisionError("tridag: rewrite equations")
    return T, Qprev, Fs, series


# --- the reference's own Fortran (compiled from its sources into _ref/) ------
_ref = None


def ref_available():
    return os.path.exists(REF_LIB)


def ref_conductionQ(nz, z, dt, Qn, Qnp1, T, ti, rhoc, emiss, Tsurf, Fgeotherm):
    """conductionq_ from conductionQ.f + tridag.for, called through the
    gfortran-convention C ABI that conduction.cpp:17-21 declares."""
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_LIB)
        _ref.conductionq_.argtypes = [_P] * 12
        _ref.conductionq_.restype = None
    T = np.array(T, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    ti = np.ascontiguousarray(np.broadcast_to(ti, (nz,)), np.float64)
    rhoc = np.ascontiguousarray(np.broadcast_to(rhoc, (nz,)), np.float64)
    a = [C.c_int(nz), None, C.c_double(dt), C.c_double(Qn), C.c_double(Qnp1), None, None, None,
         C.c_double(emiss), C.c_double(Tsurf), C.c_double(Fgeotherm), C.c_double(0.0)]
    _ref.conductionq_(C.byref(a[0]), _p(z), C.byref(a[2]), C.byref(a[3]), C.byref(a[4]), _p(T),
                      _p(ti), _p(rhoc), C.byref(a[8]), C.byref(a[9]), C.byref(a[10]),
                      C.byref(a[11]))
    return T, a[9].value, a[11].value


# --- mesh-field transfer (src/py/interpolate_mesh_field.py,
# build_mesh_transfer_tables.py), with the scipy cKDTree the scripts query ---
def transfer_field(P_src, X_src, P_tgt, mode="nearest", neighbors=5, A_src=None, A_tgt=None):
    """X_tgt (nlayers x nt) by the loops of interpolate_mesh_field.py:62-103."""
    import scipy.spatial
    X_src = np.atleast_2d(X_src)
    tree = scipy.spatial.cKDTree(P_src)
    X_tgt = np.empty((X_src.shape[0], P_tgt.shape[0]))
    for i, p in enumerate(P_tgt):
        if mode == "nearest":
            _, j = tree.query(p)
            X_tgt[:, i] = X_src[:, j]
            continue
        D, I = tree.query(p, neighbors)
        lam = D / np.sum(D)
        if mode == "weighted":
            X_tgt[:, i] = X_src[:, I] @ lam
        else:
            for layer in range(X_src.shape[0]):
                X_tgt[layer, i] = lam @ (X_src[layer, I] * A_src[I]) / A_tgt[i]
    return X_tgt


def transfer_tables(P_list, num_neighbors=5):
    """(downsample, neighbour) tables of build_mesh_transfer_tables.py:30-48."""
    import scipy.spatial
    down, neigh = {}, {}
    kd_src = scipy.spatial.cKDTree(P_list[0])
    for i in range(1, len(P_list)):
        kd_tgt = scipy.spatial.cKDTree(P_list[i])
        down[i] = np.array([kd_src.query(p)[1] for p in P_list[i]], dtype=np.uint32)
        neigh[i] = np.array([kd_tgt.query(p, num_neighbors)[1] for p in P_list[i]],
                            dtype=np.uint32)
        kd_src = kd_tgt
    return down, neigh
