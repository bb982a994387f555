"""CPU tests that pin oracle behaviours no closed form in test_cpu.py fixes:
each is chosen so that a plausible slip in the restatement fails it.

  visibility rule     a ray's closest-hit facet is inserted, not its target
                      (illum.cpp:80-94)
  albedo              rows are scaled, F(rho) = diag(rho) F(1) (illum.cpp:225-234)
  horizon frame       [b t n] with b = t x n, azimuth measured from b towards t
                      (illum.cpp:244-262, 290-297)
  fib_spiral          angles a_i = i 2 pi (phi - 1) / phi mod 2 pi (illum.cpp:468-483)
  OBJ input           first shape only, normals indexed by vertex, normal from
                      the undivided vertices, vertices / 100 (obj_util.cpp:34-92)
  tridag zero pivot   b(1) == 0 stops (tridag.for:10-12)
plus the host-side shape checks of the Python binding, which raise before any
GPU call.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OBJ = os.path.join(ROOT, "tests", "golden", "two_shapes.obj")


def _sets(bits, nf):
    return [set(np.flatnonzero(r)) for r in
            np.unpackbits(bits.view(np.uint8), bitorder="little", axis=1)[:, :nf]]


def occluder_scene():
    """Source j (z = 0, facing up), occluder k (z = 1, facing down, large) and
    target i (z = 2, facing down) on one vertical line: every ray from j to i
    passes through k."""
    V = np.array([[-0.1, -0.1, 0], [0.1, -0.1, 0], [0, 0.1, 0],       # j
                  [-1, -1, 1], [0, 1, 1], [1, -1, 1],                  # k
                  [-0.1, -0.1, 2], [0, 0.1, 2], [0.1, -0.1, 2]], float)  # i
    F = np.array([[0, 1, 2], [3, 4, 5], [6, 7, 8]], np.int32)
    return V, F


def test_visibility_inserts_the_closest_hit_not_the_target():
    V, F = occluder_scene()
    g = O.geometry(V, F)
    assert g["normal"][0][2] == 1 and g["normal"][1][2] == -1 and g["normal"][2][2] == -1
    for use_bvh in (False, True):
        bits, nrays = O.visibility(g, use_bvh=use_bvh)
        # j: rays to k and to i both stop at k; k: i is culled (behind k's
        # back), the ray to j reaches j; i: rays to k and j both stop at k
        assert _sets(bits, 3) == [{1}, {0}, {1}]
        assert nrays == 5
    cp, ri, v = O.form_factors(g, bits, albedo=1.0)
    # k shows its back to i, so F(k, i) = 0 is dropped (illum.cpp:125-127)
    assert list(cp) == [0, 1, 2, 2] and list(ri) == [1, 0]
    assert np.all(v > 0)


def test_albedo_vector_scales_rows():
    V, F = O_mesh(3, 40, seed=5)
    g = O.geometry(V, F)
    bits, _ = O.visibility(g)
    rho = np.random.default_rng(3).uniform(0.05, 0.15, len(F))
    cp1, ri1, v1 = O.form_factors(g, bits, albedo=1.0)
    cp, ri, v = O.form_factors(g, bits, albedo_vec=rho)
    assert np.array_equal(cp, cp1) and np.array_equal(ri, ri1)
    # row i (receiver) scaled by rho_i: exact, since F(1) carries a factor 1.0
    assert np.array_equal(v, v1 * rho[ri])
    cols = np.repeat(np.arange(len(F)), np.diff(cp))
    assert not np.allclose(v, v1 * rho[cols], rtol=1e-3)


def O_mesh(subdiv, ncraters, seed=0):
    from illum_amd import api
    return api.synth_ellipsoid(subdiv, ncraters=ncraters, seed=seed)


def wall_scene():
    """Facet 0: v0 = 0, v1 = e_x, v2 = e_y, so n = e_z, t = e_x and
    b = t x n = -e_y; azimuth phi points along cos(phi) b + sin(phi) t, so
    phi = pi/4 looks along (1, -1, 0)/sqrt(2).  A wall (two triangles) stands
    3 units away in that direction, 2 wide and reaching z = 5."""
    p = np.array([1 / 3, 1 / 3, 0.0])
    d = np.array([1.0, -1.0, 0.0]) / np.sqrt(2)
    u = np.array([1.0, 1.0, 0.0]) / np.sqrt(2)
    c = p + 3 * d
    z0, z1 = np.array([0, 0, -1.0]), np.array([0, 0, 5.0])
    W = [c - u + z0, c + u + z0, c + u + z1, c - u + z1]
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]] + [list(w) for w in W], float)
    F = np.array([[0, 1, 2], [3, 4, 5], [3, 5, 6]], np.int32)
    return V, F


def test_horizon_frame_azimuth_orientation():
    V, F = wall_scene()
    g = O.geometry(V, F)
    fr = g["frame"][0]
    assert np.allclose(fr[0:3], [0, -1, 0]) and np.allclose(fr[3:6], [1, 0, 0])
    nphi = 73  # 5 degree steps
    hz, _ = O.horizons(g, nphi=nphi, theta_eps=1e-3, j0=0, j1=1, use_bvh=False)
    phi = np.linspace(0, 2 * np.pi, nphi)
    dist = np.abs((phi - np.pi / 4 + np.pi) % (2 * np.pi) - np.pi)  # angle from the wall
    near, far = dist <= np.radians(12), dist >= np.radians(30)
    # towards the wall the horizon is the wall's top: atan(3 / 5) from the zenith
    assert np.all(np.abs(hz[near, 0] - np.arctan2(3.0, 5.0)) < 0.02)
    # elsewhere only the facet's own plane: within theta_eps of pi/2
    assert np.all(np.abs(hz[far, 0] - np.pi / 2) <= 1e-3 + 1e-12)
    # a mirrored frame (b = n x t) would put the wall at 7 pi / 4
    assert hz[nphi * 7 // 8, 0] > 1.5


def test_fib_spiral_angles():
    n = 100
    xy = O.fib_spiral(n)
    gr = (1 + 5 ** 0.5) / 2
    da = 2 * np.pi * (gr - 1) / gr
    i = np.arange(1, n)
    a = np.arctan2(xy[1:, 1], xy[1:, 0]) % (2 * np.pi)
    diff = (a - (i * da) % (2 * np.pi) + np.pi) % (2 * np.pi) - np.pi
    assert np.max(np.abs(diff)) < 1e-12
    np.testing.assert_allclose(np.hypot(xy[:, 0], xy[:, 1]), np.arange(n) / (n + 1), atol=1e-15)


def test_obj_first_shape_vertex_normals_and_divisor():
    V, F, VN = O.read_obj_first_shape(OBJ)
    # 8 octahedron faces and one quad (two triangles); the "second" and
    # "third" shapes are not loaded
    assert F.shape == (10, 3) and F.max() == 9
    g = O.get_objects(OBJ)
    # vertices divided by 100 in FP32
    assert np.array_equal(g["tri"].reshape(-1, 3, 3),
                          V.astype(np.float32)[F] / np.float32(100.0))
    # faces wound either way come out facing the vertex normals
    raw = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flipped = np.sum(raw * g["normal"], 1) < 0
    assert list(flipped) == [False] * 4 + [True] * 4 + [False] * 2
    assert np.all(np.sum(g["normal"] * (VN[F[:, 0]] + VN[F[:, 1]] + VN[F[:, 2]]), 1) > 0)
    # the normal is taken before dividing: the same at every divisor
    g1 = O.geometry(V, F, VN=VN, divisor=1.0)
    g3 = O.geometry(V, F, VN=VN, divisor=3.0)
    assert np.array_equal(g["normal"], g1["normal"]) and np.array_equal(g["normal"], g3["normal"])
    assert not np.array_equal(g["tri"], g3["tri"])


def zero_pivot_case():
    """A single two-layer column whose first pivot b(1) evaluates to exactly 0:
    k = 1, alpha_1 = gamma_1 = 1/2, Tr = 1, and an emissivity for which
    brad = ((2 eps) sigma) Tr^3 rounds to -0.6, so bn = 1.6 / 0.4 = 4 exactly
    and b(1) = (1 + 1/2) + 1/2 - (1/2) 4 = 0."""
    sig = 5.6704e-8
    e = -0.3 / sig
    for _ in range(200):
        brad = ((2.0 * e) * sig) * ((1.0 * 1.0) * 1.0)
        if brad == -0.6:
            break
        e = np.nextafter(e, -np.inf if brad > -0.6 else np.inf)
    assert brad == -0.6
    z = np.array([0.5, 1.0])
    return dict(nz=2, z=z, dt=1.0, Qn=10.0, Qnp1=20.0, T=np.array([1.0, 1.0]),
                ti=np.ones(2), rhoc=np.ones(2), emiss=float(e), Tsurf=1.0, Fgeotherm=0.0)


def test_zero_pivot_oracle():
    c = zero_pivot_case()
    with pytest.raises(ZeroDivisionError):
        O.conductionQ(**c)
    # one ulp of emissivity away the step goes through
    c2 = dict(c, emiss=float(np.nextafter(c["emiss"], 0)))
    O.conductionQ(**c2)


def test_binding_shape_checks_raise_before_any_gpu_call():
    from illum_amd import _lib, api
    # a T0 profile of the wrong shape (thermal.hpp:53-57)
    with pytest.raises(_lib.IllumError) as e:
        api.ThermalModel(10, np.linspace(0.01, 0.5, 5), 200.0, 1.2e6, np.zeros((5, 10)))
    assert e.value.code == _lib.ILLUM_ERR_SHAPE
    th = object.__new__(api.ThermalModel)
    th.nf, th.h = 10, None
    for bad in (np.zeros(9), np.zeros((2, 9))):
        with pytest.raises(_lib.IllumError) as e:
            th.step(1.0, bad)
        assert e.value.code == _lib.ILLUM_ERR_SHAPE
    with pytest.raises(_lib.IllumError):
        th.run(1.0, np.zeros((3, 11)))
    # CSR arrays shorter than rowptr says
    with pytest.raises(_lib.IllumError):
        api.SparseMatrix.from_csr(None, 3, [0, 1, 2, 3], [0, 1], [0.1, 0.2, 0.3])
    with pytest.raises(_lib.IllumError):
        api.SparseMatrix.from_csr(None, 3, [0, 1, 3], [0, 1, 2], [0.1, 0.2, 0.3])
    # horizons that do not match the mesh
    ic = object.__new__(api.IllumContext)
    ic.num_faces, ic.horizons = 5, np.zeros((9, 4))
    with pytest.raises(_lib.IllumError):
        ic.get_direct_radiosity(np.array([1e8, 0, 0]), O.fib_spiral(100))


def test_transfer_files_use_the_scripts_c_order_convention(tmp_path):
    # src/py/arma.py writes the header, "n" for a vector or "rows cols", then
    # the values in C order; the transfer tools read and write the same bytes
    from illum_amd import transfer
    X = np.arange(6, dtype=np.float64).reshape(2, 3)
    transfer.arma_tofile(X, tmp_path / "x.bin")
    raw = open(tmp_path / "x.bin", "rb").read()
    assert raw.startswith(b"ARMA_MAT_BIN_FN008\n2 3\n")
    assert np.array_equal(np.frombuffer(raw[23:], np.float64), X.ravel())
    assert np.array_equal(transfer.arma_fromfile(tmp_path / "x.bin"), X)
    u = np.array([3, 1, 2], np.uint32)
    transfer.arma_tofile(u, tmp_path / "u.bin")
    assert open(tmp_path / "u.bin", "rb").read().startswith(b"ARMA_MAT_BIN_IU004\n3\n")
    assert np.array_equal(transfer.arma_fromfile(tmp_path / "u.bin", np.uint32), u)


def _two_leaf_tree(tri9):
    """A tree in the exported layout over nf <= 254 triangles: one inner node
    whose children are leaves of the first and second half; threaded nodes
    root, left leaf, right leaf."""
    nf = tri9.shape[0]
    h = nf // 2
    t = tri9.reshape(nf, 3, 3)
    pad = np.float32(1e-4)

    def box(a, b):
        return t[a:b].reshape(-1, 3).min(0) - pad, t[a:b].reshape(-1, 3).max(0) + pad

    def fbits(i):
        return np.array([i], np.int32).view(np.float32)[0]

    info = lambda first, cnt: (cnt << 24) | first
    (l0, l1), (r0, r1), (b0, b1) = box(0, h), box(h, nf), box(0, nf)
    wide = np.zeros((1, 16), np.float32)
    wide[0, 0:3], wide[0, 3], wide[0, 4:7] = l0, fbits(~info(0, h)), l1
    wide[0, 8:11], wide[0, 11], wide[0, 12:15] = r0, fbits(~info(h, nf - h)), r1
    nodes = np.zeros((3, 8), np.float32)
    nodes[0, 0:3], nodes[0, 3], nodes[0, 4:7], nodes[0, 7] = b0, fbits(-1), b1, fbits(-1)
    nodes[1, 0:3], nodes[1, 3], nodes[1, 4:7], nodes[1, 7] = l0, fbits(2), l1, fbits(info(0, h))
    nodes[2, 0:3], nodes[2, 3], nodes[2, 4:7], nodes[2, 7] = r0, fbits(-1), r1, fbits(info(h, nf - h))
    tris = np.zeros((nf, 12), np.float32)
    tris[:, 0:3], tris[:, 4:7], tris[:, 8:11] = t[:, 0], t[:, 1], t[:, 2]
    tris[:, 3] = np.arange(nf, dtype=np.int32).view(np.float32)
    return dict(nodes=nodes, wide=wide, tris=tris, roots=np.array([[0, 0, -1]], np.int32))


def test_walk_work_on_a_two_leaf_tree():
    # The per-ray walk that prices the visibility kernel's roofline finds the
    # same closest hits as the brute-force visibility, casts the same rays,
    # and visits the one inner node once per walk (2 box tests); with no stack
    # entries every walk that enters both leaves is answered by the threaded
    # walk instead, still with the same hits.
    V, F = O_mesh(1, 6, seed=2)
    g = O.geometry(V, F)
    tree = _two_leaf_tree(g["tri"])
    bits, nrays = O.visibility(g, use_bvh=False)
    w, wbits = O.walk_work(g, tree, with_bits=True)
    assert np.array_equal(wbits, bits) and w["rays"] == nrays > 0
    # every ray's first walk is bounded at t <= 1.00098 and hits its target
    # facet's plane at about t = 1 (closed mesh), so no second walk runs
    assert w["box_tests"] == 2 * nrays
    assert nrays <= w["tri_tests"] <= len(F) * nrays
    w0, b0 = O.walk_work(g, tree, cap=0, with_bits=True)
    assert np.array_equal(b0, bits) and w0["rays"] == nrays
    assert w0["box_tests"] > w["box_tests"]
