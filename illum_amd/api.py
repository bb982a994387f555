"""Host-side mirror of the reference interface for the form-factor / horizon /
radiosity / thermal path, on top of the C ABI (include/illum_hip.h).

Reference names kept:
  IllumContext.compute_F            illum_context::compute_F       (illum.hpp:18-20)
  IllumContext.make_horizons        illum_context::make_horizons   (illum.hpp:22-23)
  IllumContext.save/load_horizons   (illum.hpp:25-27)
  IllumContext.get_direct_radiosity (illum.hpp:29-30)
  IllumContext.get_num_faces        (illum.hpp:32-34)
  IllumContext.set_horizon_obj_file (illum.hpp:36-38)
  fib_spiral                        (illum.hpp:50)
  ThermalModel(nfaces, depths, ti, rhoc, T0).step(dt, Q) / get_radiosity()
                                    thermal_model (thermal.hpp:17-105)
  conductionQ                       conduction.hpp:16-28
  gauss_seidel-style solve          radiosity_solve (illum_cli.cpp:425-440)

Arrays are numpy (host) by default; DeviceArray (HBM allocated by the library)
is passed as a device pointer with no host staging.  Every compute call runs
on the GPU through libillum_hip; nothing here computes results on the CPU.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import call

_PP = C.c_void_p


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


def _is_dev(a):
    return a is not None and hasattr(a, "is_cuda") and a.is_cuda


def _shape_error(msg):
    return _lib.IllumError(_lib.ILLUM_ERR_SHAPE, msg)


def _nbytes(a):
    return a.nbytes if hasattr(a, "nbytes") else int(np.prod(a.shape)) * np.dtype(a.dtype).itemsize


class Context:
    """One GPU + one HIP stream (illum_ctx).

    Objects made on a context (meshes, BVHs, matrices, thermal models, device
    arrays) must be closed before it.  One closed after its context (the
    interpreter's shutdown finalizes objects in any order) only drops its
    handle: destroying it would touch the freed context, so its buffers are
    left to the end of the process."""

    def __init__(self, device=0):
        h = C.c_void_p()
        call("illum_ctx_create", device, C.byref(h))
        self.h = h

    def close(self):
        if self.h:
            call("illum_ctx_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        call("illum_ctx_synchronize", self.h)

    def set_profiling(self, on):
        call("illum_ctx_set_profiling", self.h, 1 if on else 0)

    def reset_profile(self):
        call("illum_ctx_reset_profile", self.h)

    def tuning(self):
        """The autotuner's choices on this context: [(key, choice, best_ms, trials)]."""
        n = C.c_int()
        call("illum_ctx_tuning", self.h, None, 0, C.byref(n))
        buf = C.create_string_buffer(n.value + 1)
        call("illum_ctx_tuning", self.h, buf, n.value + 1, C.byref(n))
        out = []
        for line in buf.value.decode().splitlines():
            f = line.split()
            out.append((f[0], int(f[1]), float(f[2]), f[3:]))
        return out

    def profile(self):
        """{stage: (total_ms, launches)} of hipEvent-timed stages."""
        cap = 64
        ms = np.zeros(cap)
        cnt = np.zeros(cap, dtype=np.int64)
        names = C.create_string_buffer(8192)
        n = C.c_int()
        call("illum_ctx_get_profile", self.h, cap, _ptr(ms), _ptr(cnt), names, 8192, C.byref(n))
        keys = names.raw.split(b"\0")[: n.value]
        return {k.decode(): (float(ms[i]), int(cnt[i])) for i, k in enumerate(keys)}


class DeviceArray:
    """A buffer in the context GPU's HBM (illum_dev_alloc), usable wherever the
    API takes a device pointer.  Row-major numpy-style shape and dtype."""

    is_cuda = True

    def __init__(self, ctx, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(shape) if np.ndim(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        call("illum_dev_alloc", ctx.h, self.nbytes, C.byref(p))
        self.ptr = p

    @classmethod
    def from_host(cls, ctx, arr):
        a = np.ascontiguousarray(arr)
        d = cls(ctx, a.shape, a.dtype)
        call("illum_dev_copy", ctx.h, d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, 0)
        return d

    def to_host(self):
        a = np.empty(self.shape, self.dtype)
        call("illum_dev_copy", self.ctx.h, a.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, 1)
        return a

    def to_host_flat(self, start, count):
        """Elements [start, start + count) of the flat (row-major) buffer."""
        if start < 0 or count < 0 or (start + count) * self.dtype.itemsize > self.nbytes:
            raise _shape_error("slice outside the device array")
        a = np.empty(count, self.dtype)
        src = C.c_void_p(self.ptr.value + start * self.dtype.itemsize)
        call("illum_dev_copy", self.ctx.h, a.ctypes.data_as(C.c_void_p), src, a.nbytes, 1)
        return a

    def data_ptr(self):
        return self.ptr.value

    def dim(self):
        return len(self.shape)

    def close(self):
        if self.ptr and self.ctx.h:
            call("illum_dev_free", self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def synth_ellipsoid(subdiv, axes=(13.0, 11.0, 9.0), ncraters=40, depth=(0.05, 0.3),
                    width=(0.2, 1.5), seed=0):
    """Seeded cratered triaxial ellipsoid: returns (V [nv,3] f64, F [nf,3] i32)."""
    nv = C.c_int64()
    nf = C.c_int64()
    call("illum_synth_sizes", subdiv, C.byref(nv), C.byref(nf))
    V = np.zeros((nv.value, 3))
    F = np.zeros((nf.value, 3), dtype=np.int32)
    call("illum_synth_ellipsoid", subdiv, axes[0], axes[1], axes[2], ncraters, depth[0],
         depth[1], width[0], width[1], seed, _ptr(V), _ptr(F))
    return V, F


def concat_bodies(bodies):
    """[(V, F), ...] -> (V, F, offsets): one mesh of several bodies, facet
    offsets for Mesh(bodies=...)."""
    Vs, Fs, off, vbase = [], [], [0], 0
    for V, F in bodies:
        Vs.append(np.asarray(V, np.float64))
        Fs.append(np.asarray(F, np.int32) + vbase)
        vbase += len(V)
        off.append(off[-1] + len(F))
    return np.concatenate(Vs), np.ascontiguousarray(np.concatenate(Fs)), np.asarray(off, np.int64)


class Mesh:
    """bodies: optional facet offsets (nbody + 1, from 0 to N): facets
    [bodies[b], bodies[b+1]) form an independent body (illum_mesh_set_bodies)."""

    def __init__(self, ctx, V=None, F=None, VN=None, vertex_divisor=1.0, obj_path=None,
                 bodies=None):
        self.ctx = ctx
        h = C.c_void_p()
        if obj_path is not None:
            call("illum_mesh_from_obj", ctx.h, str(obj_path).encode(), vertex_divisor, C.byref(h))
        else:
            V = np.ascontiguousarray(V, dtype=np.float64)
            F = np.ascontiguousarray(F, dtype=np.int32)
            VN = None if VN is None else np.ascontiguousarray(VN, dtype=np.float64)
            call("illum_mesh_create", ctx.h, _ptr(V), V.shape[0], _ptr(F), F.shape[0], _ptr(VN),
                 vertex_divisor, C.byref(h))
        self.h = h
        n = C.c_int64()
        call("illum_mesh_num_faces", h, C.byref(n))
        self.nf = n.value
        self.nbody = 1
        if bodies is not None:
            off = np.ascontiguousarray(bodies, np.int64)
            call("illum_mesh_set_bodies", h, _ptr(off), len(off) - 1)
            self.nbody = len(off) - 1

    def close(self):
        if self.h and self.ctx.h:
            call("illum_mesh_destroy", self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geometry(self):
        nf = self.nf
        cen = np.zeros((nf, 3), np.float32)
        nrm = np.zeros((nf, 3), np.float32)
        rad = np.zeros(nf, np.float32)
        area = np.zeros(nf)
        frame = np.zeros((nf, 9))
        call("illum_mesh_geometry", self.h, _ptr(cen), _ptr(nrm), _ptr(rad), _ptr(area), _ptr(frame))
        return dict(centroid=cen, normal=nrm, radius=rad, area=area, frame=frame)

    def triangles(self):
        t = np.zeros((self.nf, 9), np.float32)
        call("illum_mesh_triangles", self.h, _ptr(t))
        return t


class BVH:
    def __init__(self, mesh, leaf_size=4):
        self.mesh = mesh
        h = C.c_void_p()
        call("illum_bvh_build", mesh.h, leaf_size, C.byref(h))
        self.h = h

    def close(self):
        if self.h and self.mesh.h and self.mesh.ctx.h:
            call("illum_bvh_destroy", self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        n = C.c_int64()
        d = C.c_int()
        l = C.c_int()
        call("illum_bvh_info", self.h, C.byref(n), C.byref(d), C.byref(l))
        return dict(nodes=n.value, depth=d.value, leaf_size=l.value)

    def export(self):
        """The tree as the walks read it (illum_bvh_export): dict of nodes
        (threaded, n x 8), wide (two-wide, n x 16), tris (n x 12) float32 and
        roots (bodies x 3: threaded, two-wide, four-wide) int32."""
        sz = np.zeros(4, np.int64)
        call("illum_bvh_export", self.h, _ptr(sz), None, None, None, None)
        out = dict(nodes=np.zeros((sz[0], 8), np.float32), wide=np.zeros((sz[1], 16), np.float32),
                   tris=np.zeros((sz[2], 12), np.float32), roots=np.zeros((sz[3], 3), np.int32))
        call("illum_bvh_export", self.h, _ptr(sz), _ptr(out["nodes"]), _ptr(out["wide"]),
             _ptr(out["tris"]), _ptr(out["roots"]))
        return out

    def tuning(self):
        """[(leaf_size, layout, ms)] of the autotuning trials (empty unless
        leaf_size was 0); layout 2 = two-wide, 3 = wavefront packet over the
        two-wide tree, 4 = four-wide, 5 = wavefront packet over the four-wide
        tree."""
        leaves = np.zeros(32, np.int32)
        layouts = np.zeros(32, np.int32)
        ms = np.zeros(32)
        n = C.c_int()
        call("illum_bvh_tuning", self.h, 32, _ptr(leaves), _ptr(layouts), _ptr(ms), C.byref(n))
        return [(int(leaves[k]), int(layouts[k]), float(ms[k])) for k in range(n.value)]

    def closest_hit(self, orig, dirs):
        orig = np.ascontiguousarray(orig, np.float32)
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = orig.shape[0]
        hit = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        call("illum_bvh_closest_hit", self.h, _ptr(orig), _ptr(dirs), n, _ptr(hit), _ptr(t))
        return hit, t


class SparseMatrix:
    """N x N form-factor matrix held on the GPU in CSC and CSR form."""

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h
        n = C.c_int64()
        nnz = C.c_int64()
        call("illum_sparse_info", h, C.byref(n), C.byref(nnz))
        self.n = n.value
        self.nnz = nnz.value

    @classmethod
    def from_csr(cls, ctx, n, rowptr, colidx, val):
        rowptr = np.ascontiguousarray(rowptr, np.int64)
        colidx = np.ascontiguousarray(colidx, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        # illum_sparse_create_csr reads rowptr[0..n] and rowptr[n] entries of
        # colidx and val
        if rowptr.shape != (n + 1,):
            raise _shape_error("rowptr must hold n + 1 offsets")
        nnz = int(rowptr[-1])
        if colidx.shape != (nnz,) or val.shape != (nnz,):
            raise _shape_error("colidx and val must hold rowptr[n] entries")
        h = C.c_void_p()
        call("illum_sparse_create_csr", ctx.h, n, int(rowptr[-1]), _ptr(rowptr), _ptr(colidx),
             _ptr(val), C.byref(h))
        return cls(ctx, h)

    def close(self):
        if self.h and self.ctx.h:
            call("illum_sparse_destroy", self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def csc(self):
        cp = np.zeros(self.n + 1, np.int64)
        ri = np.zeros(self.nnz, np.int32)
        v = np.zeros(self.nnz)
        call("illum_sparse_copy_csc", self.h, _ptr(cp), _ptr(ri), _ptr(v))
        return cp, ri, v

    def csr(self):
        rp = np.zeros(self.n + 1, np.int64)
        ci = np.zeros(self.nnz, np.int32)
        v = np.zeros(self.nnz)
        call("illum_sparse_copy_csr", self.h, _ptr(rp), _ptr(ci), _ptr(v))
        return rp, ci, v

    def visibility_lists(self):
        """Raw visibility sets before the zero drop (assembled with
        keep_visibility): (ptr, idx) with source j's hits idx[ptr[j]:ptr[j+1]]."""
        nv = C.c_int64()
        call("illum_sparse_visibility_nnz", self.h, C.byref(nv))
        ptr = np.zeros(self.n + 1, np.int64)
        idx = np.zeros(nv.value, np.int32)
        call("illum_sparse_copy_visibility_lists", self.h, _ptr(ptr), _ptr(idx))
        return ptr, idx

    def visibility(self):
        """Raw visibility bit rows (n x ceil(n/32) uint32) before the zero drop."""
        ptr, idx = self.visibility_lists()
        words = (self.n + 31) // 32
        bits = np.zeros((self.n, words), np.uint32)
        rows = np.repeat(np.arange(self.n), np.diff(ptr))
        np.bitwise_or.at(bits, (rows, idx >> 5), (np.uint32(1) << (idx & 31).astype(np.uint32)))
        return bits

    def rays(self):
        """Visibility rays traced when the matrix was assembled."""
        n = C.c_int64()
        call("illum_sparse_rays", self.h, C.byref(n))
        return n.value


    def to_scipy(self):
        import scipy.sparse as sps
        cp, ri, v = self.csc()
        return sps.csc_matrix((v, ri, cp), shape=(self.n, self.n))


def fib_spiral(n):
    """fib_spiral(xy, n) (illum.cpp:468-483): n x 2 disk sample points."""
    xy = np.zeros((2, n))  # column-major n x 2 == row-major 2 x n
    call("illum_fib_spiral", _ptr(xy), n)
    return xy.T.copy()


class IllumContext:
    """Mirror of illum_context (illum.hpp:6-48) over the GPU library."""

    def __init__(self, obj_path=None, V=None, F=None, VN=None, vertex_divisor=None,
                 leaf_size=4, ctx=None, bodies=None):
        self.ctx = ctx or default_context()
        if obj_path is not None:
            div = 100.0 if vertex_divisor is None else vertex_divisor
            self.mesh = Mesh(self.ctx, obj_path=obj_path, vertex_divisor=div, bodies=bodies)
        else:
            div = 1.0 if vertex_divisor is None else vertex_divisor
            self.mesh = Mesh(self.ctx, V=V, F=F, VN=VN, vertex_divisor=div, bodies=bodies)
        self.bvh = BVH(self.mesh, leaf_size)
        self.num_faces = self.mesh.nf
        self.horizons = None
        self._horizon_ctx = None
        self.leaf_size = leaf_size

    def get_num_faces(self):
        return self.num_faces

    def set_horizon_obj_file(self, path, vertex_divisor=100.0):
        occ = Mesh(self.ctx, obj_path=path, vertex_divisor=vertex_divisor)
        self._horizon_ctx = (occ, BVH(occ, self.leaf_size))

    def compute_F(self, albedo, offset=1e-5, keep_visibility=False):
        """albedo: scalar, per-facet array, or a path to an Armadillo vector.
        keep_visibility: also keep the raw visibility sets (testing aid)."""
        if isinstance(albedo, str):
            from .armaio import load_mat
            albedo = load_mat(albedo).ravel()
        h = C.c_void_p()
        flags = _lib.ILLUM_FF_KEEP_VISIBILITY if keep_visibility else 0
        if np.isscalar(albedo):
            call("illum_form_factors_ex", self.bvh.h, float(albedo), None, offset, flags,
                 C.byref(h))
        else:
            a = np.ascontiguousarray(albedo, np.float64)
            if a.shape != (self.num_faces,):
                raise _shape_error("albedo vector must have one entry per facet")
            call("illum_form_factors_ex", self.bvh.h, 0.0, _ptr(a), offset, flags, C.byref(h))
        return SparseMatrix(self.ctx, h)

    def render(self, data=None, mode="orthographic", width=512, height=512, az=0.0, el=0.0,
               minval=None, maxval=None):
        """The renderer's ray-cast (illum_renderer_cli.cpp:174-261): returns
        (facet image, value image), height x width, row 0 at the top; facet -1
        and value NaN where no facet is hit; value is None without data."""
        m = {"orthographic": 0, "equal_area_cylindrical": 1}[mode]
        facet = np.zeros((height, width), np.int32)
        value = None
        d = None
        if data is not None:
            d = np.ascontiguousarray(data, np.float64).ravel()
            if d.shape != (self.num_faces,):
                raise ValueError("input data and mesh have incompatible sizes")
            value = np.zeros((height, width))
        lo = -np.inf if minval is None else float(minval)
        hi = np.inf if maxval is None else float(maxval)
        call("illum_render", self.bvh.h, m, width, height, float(az), float(el), _ptr(d), lo, hi,
             _ptr(facet), _ptr(value), 0)
        return facet, value

    def form_factor_work(self, offset=1e-5):
        """Work of one visibility pass (separate instrumented launch): rays, pair
        (facing) tests after the cluster pass, box and triangle tests per active
        lane, and for the wavefront walks their walks, node visits and triangle
        tests per wavefront and the walks with a fixed axis permutation."""
        out = np.zeros(8, np.int64)
        call("illum_form_factors_work_ex", self.bvh.h, offset, _ptr(out), 8)
        return dict(zip(("rays", "pair_tests", "box_tests", "tri_tests", "wave_walks",
                         "wave_node_visits", "wave_tri_tests", "wave_walks_fixed_perm"),
                        map(int, out)))

    def make_horizons(self, nphi=361, theta_eps=1e-3, offset=1e-5, out=None):
        occ = self._horizon_ctx[1].h if self._horizon_ctx else None
        if out is not None and _is_dev(out):
            call("illum_horizons", self.bvh.h, occ, nphi, theta_eps, offset, _ptr(out),
                 _lib.ILLUM_OUT_ON_DEVICE)
            self.horizons = out
            return out
        hz = np.zeros((self.num_faces, nphi))  # row f = column f of the nphi x N matrix
        call("illum_horizons", self.bvh.h, occ, nphi, theta_eps, offset, _ptr(hz), 0)
        self.horizons = hz.T
        return self.horizons

    def save_horizons(self, path):
        from .armaio import save_mat
        save_mat(np.asfortranarray(self.horizons), path)

    def load_horizons(self, path):
        from .armaio import load_mat
        self.horizons = load_mat(path)

    def get_direct_radiosity(self, sun_position, disk_xy, per_body=False):
        """sun_position: (3,) or (3, S) km; returns E of shape (N,) or (N, S).
        per_body: a multi-body mesh's bodies each have their own S suns,
        sun_position is (3, S * nbody) with body b's in columns b S .. b S + S - 1."""
        sun = np.asarray(sun_position, np.float64)
        single = sun.ndim == 1
        sun = np.ascontiguousarray(sun.reshape(3, -1).T)  # S x 3 row-major == 3 x S col-major
        S = sun.shape[0]
        flags = 0
        if per_body:
            nb = self.mesh.nbody
            if S % nb:
                raise _shape_error("per-body suns must be S x nbody columns")
            S //= nb
            single = False
            flags = _lib.ILLUM_SUN_PER_BODY
        disk = np.asfortranarray(np.asarray(disk_xy, np.float64))
        if disk.ndim != 2 or disk.shape[1] != 2:
            raise _shape_error("disk_xy must be n x 2")
        if self.horizons is None:
            raise _lib.IllumError(_lib.ILLUM_ERR_ARG, "no horizons: call make_horizons first")
        hz = np.asfortranarray(self.horizons)
        if hz.ndim != 2 or hz.shape[1] != self.num_faces or hz.shape[0] < 2:
            raise _shape_error("horizons must be nphi x N")
        nphi = hz.shape[0]
        E = np.zeros((S, self.num_faces))
        call("illum_direct_radiosity", self.mesh.h, hz.ctypes.data_as(C.c_void_p), nphi,
             _ptr(sun), S, disk.ctypes.data_as(C.c_void_p), disk.shape[0], _ptr(E), flags)
        return E[0] if single else E.T


def radiosity_solve(F, E, method="mixed", rel_tol=1e-8, max_iter=200, B=None):
    """Solve (I - F) B = E for E of shape (N,) or (N, S).  Returns (B, resid, iters).

    Device arrays (DeviceArray) hold the C ABI's column-major N x S block,
    i.e. shape (N,) for one right-hand side or (S, N) row-major for S of
    them; B must have E's shape and dtype float64."""
    m = {"fp64": _lib.ILLUM_RAD_JACOBI_FP64, "mixed": _lib.ILLUM_RAD_MIXED}[method]
    resid = np.zeros(max_iter + 1)
    iters = C.c_int()
    if _is_dev(E):
        n = F.n
        if E.dim() == 1:
            S = 1
            if E.shape != (n,):
                raise _shape_error("device E must have shape (N,) or (S, N)")
        elif E.dim() == 2 and E.shape[1] == n:
            S = E.shape[0]
        else:
            raise _shape_error("device E must have shape (N,) or (S, N)")
        if B is None or not _is_dev(B):
            raise _lib.IllumError(_lib.ILLUM_ERR_ARG, "a device E needs a device B")
        for a in (E, B):
            if np.dtype(a.dtype) != np.float64 or _nbytes(a) != 8 * n * S:
                raise _shape_error("E and B must each hold N x S doubles")
        call("illum_radiosity_solve", F.h, _ptr(E), S, m, rel_tol, max_iter, _ptr(B),
             _ptr(resid), C.byref(iters),
             _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_OUT_ON_DEVICE | _lib.ILLUM_ASYNC)
        out = B
    else:
        E2 = np.asarray(E, np.float64)
        if E2.ndim not in (1, 2) or E2.shape[0] != F.n:
            raise _shape_error("E must have shape (N,) or (N, S)")
        single = E2.ndim == 1
        Ecm = np.ascontiguousarray(E2.reshape(E2.shape[0], -1).T)  # S x N == N x S col-major
        S = Ecm.shape[0]
        Bcm = np.zeros_like(Ecm)
        call("illum_radiosity_solve", F.h, _ptr(Ecm), S, m, rel_tol, max_iter, _ptr(Bcm),
             _ptr(resid), C.byref(iters), 0)
        out = Bcm[0] if single else Bcm.T
    # one relative residual per product (the solve's stop decision after it)
    return out, resid[:iters.value], iters.value


class ThermalModel:
    """Mirror of thermal_model (thermal.hpp:7-105) for many facets on the GPU.

    depths: one grid (list) for all facets, or a list of grids with
    grid_of_facet selecting one per facet.  thermal_inertia: scalar or
    per-layer list (per grid); ti_scale: optional per-facet multiplier.
    """

    def __init__(self, nfaces, depths, thermal_inertia, rhoc, init_temp, emissivity=0.999,
                 precision="fp64", grid_of_facet=None, ti_scale=None, ctx=None):
        grids = depths if (len(depths) and np.ndim(depths[0]) == 1) else [depths]
        zs, tis, rcs, off = [], [], [], [0]
        for g in grids:
            g = np.asarray(g, np.float64)
            zs.append(g)
            ti = np.broadcast_to(np.asarray(thermal_inertia, np.float64), g.shape)
            rc = np.broadcast_to(np.asarray(rhoc, np.float64), g.shape)
            tis.append(ti)
            rcs.append(rc)
            off.append(off[-1] + len(g))
        self.z = np.concatenate(zs)
        self.goff = np.asarray(off, np.int64)
        ti = np.ascontiguousarray(np.concatenate(tis))
        rc = np.ascontiguousarray(np.concatenate(rcs))
        self.nz_max = max(len(g) for g in zs)
        self.nf = nfaces
        gof = None if grid_of_facet is None else np.ascontiguousarray(grid_of_facet, np.int32)
        tsc = None if ti_scale is None else np.ascontiguousarray(ti_scale, np.float64)
        prec = {"fp64": _lib.ILLUM_THERM_FP64, "mixed": _lib.ILLUM_THERM_MIXED}[precision]
        T0 = None
        T0s = 0.0
        if isinstance(init_temp, str):
            from .armaio import load_mat
            init_temp = load_mat(init_temp)
        if np.isscalar(init_temp):
            T0s = float(init_temp)
        else:
            T0 = np.asfortranarray(np.asarray(init_temp, np.float64))
            if T0.shape != (self.nz_max + 1, nfaces):
                raise _lib.IllumError(_lib.ILLUM_ERR_SHAPE,
                                      "initial temperature profile is an incompatible size")
        if gof is not None and gof.shape != (nfaces,):
            raise _shape_error("grid_of_facet must hold one grid per facet")
        if tsc is not None and tsc.shape != (nfaces,):
            raise _shape_error("ti_scale must hold one factor per facet")
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        call("illum_thermal_create", self.ctx.h, nfaces, len(grids), _ptr(self.goff),
             _ptr(self.z), _ptr(ti), _ptr(rc), _ptr(gof), _ptr(tsc), emissivity, T0s,
             None if T0 is None else T0.ctypes.data_as(C.c_void_p), prec, C.byref(h))
        self.h = h
        self.t = 0.0

    def close(self):
        if self.h and self.ctx.h:
            call("illum_thermal_destroy", self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_q(self, Q, nsteps=None):
        want = (self.nf,) if nsteps is None else (nsteps, self.nf)
        if _is_dev(Q):
            if np.dtype(Q.dtype) != np.float64 or _nbytes(Q) != 8 * int(np.prod(want)):
                raise _shape_error(f"Q must hold {want} doubles")
        elif np.shape(Q) != want:
            raise _shape_error(f"Q must have shape {want}")

    def step(self, dt, Q):
        self._check_q(Q)
        if _is_dev(Q):
            # device Q: queued without waiting (DeviceArray reads are ordered after it)
            call("illum_thermal_step", self.h, dt, _ptr(Q),
                 _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_ASYNC)
        else:
            q = np.ascontiguousarray(Q, np.float64)
            call("illum_thermal_step", self.h, dt, _ptr(q), 0)
        self.t += dt

    def run(self, dt, Qseries):
        """nsteps consecutive steps; Qseries has shape (nsteps, N)."""
        if len(np.shape(Qseries)) != 2:
            raise _shape_error("Qseries must have shape (nsteps, N)")
        self._check_q(Qseries, np.shape(Qseries)[0])
        if _is_dev(Qseries):
            n = Qseries.shape[0]
            call("illum_thermal_run", self.h, dt, _ptr(Qseries), n,
                 _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_ASYNC)
        else:
            q = np.ascontiguousarray(Qseries, np.float64)
            n = q.shape[0]
            call("illum_thermal_run", self.h, dt, _ptr(q), n, 0)
        self.t += dt * n

    def run_held(self, dt, Q, nsteps):
        """nsteps consecutive steps under one flux Q of shape (N,), held for
        all of them (one launch sequence; the same steps as nsteps calls of
        step(dt, Q))."""
        self._check_q(Q)
        if nsteps < 1:
            raise _lib.IllumError(_lib.ILLUM_ERR_ARG, "nsteps must be >= 1")
        flags = _lib.ILLUM_Q_HELD
        if _is_dev(Q):
            call("illum_thermal_run", self.h, dt, _ptr(Q), nsteps,
                 flags | _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_ASYNC)
        else:
            q = np.ascontiguousarray(Q, np.float64)
            call("illum_thermal_run", self.h, dt, _ptr(q), nsteps, flags)
        self.t += dt * nsteps

    @property
    def T(self):
        T = np.zeros((self.nf, self.nz_max + 1))
        call("illum_thermal_get", self.h, _ptr(T), None, None, 0)
        return T.T

    def surface(self):
        Ts = np.zeros(self.nf)
        Fs = np.zeros(self.nf)
        call("illum_thermal_get", self.h, None, _ptr(Ts), _ptr(Fs), 0)
        return Ts, Fs

    def get_radiosity(self, out=None, accumulate=False):
        """-Fsurf (thermal.hpp:61-63).  With `out` (host array or DeviceArray of
        N doubles) the result is written, or added when `accumulate`, in place."""
        flags = _lib.ILLUM_ACCUMULATE if accumulate else 0
        if out is None:
            if accumulate:
                raise ValueError("accumulate needs an output array")
            out = np.zeros(self.nf)
        elif _is_dev(out):
            flags |= _lib.ILLUM_OUT_ON_DEVICE
        elif not (isinstance(out, np.ndarray) and out.dtype == np.float64 and
                  out.flags.c_contiguous and out.size == self.nf):
            raise ValueError("out must be a contiguous float64 array of N values")
        call("illum_thermal_get_radiosity", self.h, _ptr(out), flags)
        return out


class TimeLoop:
    """The coupled radiosity + thermal time loop (illum_cli.cpp:403-456) kept
    on the device and replayed from captured HIP graphs (illum_timeloop_*).

    F: SparseMatrix, illum: IllumContext (its mesh), thermal: ThermalModel;
    horizons: DeviceArray of shape (N, nphi) (the nphi x N map column-major);
    disk: DeviceArray of shape (2, ndisk) (the ndisk x 2 spiral column-major);
    products: the solve's product budget per step (0 = chosen from the first
    step); eager: queue kernels instead of replaying graphs."""

    def __init__(self, F, illum, thermal, horizons, disk, method="mixed", rel_tol=1e-8,
                 products=0, eager=False):
        m = {"fp64": _lib.ILLUM_RAD_JACOBI_FP64, "mixed": _lib.ILLUM_RAD_MIXED}[method]
        if horizons.shape[0] != illum.num_faces or len(horizons.shape) != 2:
            raise _shape_error("horizons must be a DeviceArray of shape (N, nphi)")
        if len(disk.shape) != 2 or disk.shape[0] != 2:
            raise _shape_error("disk must be a DeviceArray of shape (2, ndisk)")
        self.ctx = illum.ctx
        self.nf = illum.num_faces
        self._keep = (F, illum, thermal, horizons, disk)
        h = C.c_void_p()
        call("illum_timeloop_create", F.h, illum.mesh.h, thermal.h, _ptr(horizons),
             horizons.shape[1], _ptr(disk), disk.shape[1], m, rel_tol, products,
             _lib.ILLUM_TL_EAGER if eager else 0, C.byref(h))
        self.h = h

    def run(self, suns, dt):
        """suns: (3, nsteps) km on the host, or a DeviceArray of shape (nsteps, 3)."""
        if _is_dev(suns):
            n = suns.shape[0]
            call("illum_timeloop_run", self.h, _ptr(suns), n, dt, _lib.ILLUM_IN_ON_DEVICE)
        else:
            sc = np.ascontiguousarray(np.asarray(suns, np.float64).reshape(3, -1).T)
            call("illum_timeloop_run", self.h, _ptr(sc), sc.shape[0], dt, 0)

    def stats(self):
        st = C.c_int64()
        p, mp, un = C.c_int(), C.c_int(), C.c_int()
        r = C.c_double()
        call("illum_timeloop_stats", self.h, C.byref(st), C.byref(p), C.byref(mp), C.byref(un),
             C.byref(r))
        return dict(steps=st.value, products=p.value, max_products=mp.value,
                    unconverged=un.value, max_resid=r.value)

    def get(self):
        """(E, B) of the last step."""
        E = np.zeros(self.nf)
        B = np.zeros(self.nf)
        call("illum_timeloop_get", self.h, _ptr(E), _ptr(B), 0)
        return E, B

    def close(self):
        if self.h and self.ctx.h:
            call("illum_timeloop_destroy", self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def conductionQ(nz, z, dt, Qn, Qnp1, T, ti, rhoc, emiss, Tsurf, Fgeotherm):
    """conductionQ (conduction.hpp:16-28) on the GPU; returns (T, Tsurf, Fsurf)."""
    Tc = np.array(T, np.float64)
    ts = C.c_double(Tsurf)
    fs = C.c_double(0.0)
    z = np.ascontiguousarray(z, np.float64)
    ti = np.ascontiguousarray(np.broadcast_to(ti, (nz,)), np.float64)
    rhoc = np.ascontiguousarray(np.broadcast_to(rhoc, (nz,)), np.float64)
    call("illum_conductionQ", nz, _ptr(z), dt, Qn, Qnp1, _ptr(Tc), _ptr(ti), _ptr(rhoc), emiss,
         C.byref(ts), Fgeotherm, C.byref(fs))
    return Tc, ts.value, fs.value
