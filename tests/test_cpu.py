"""CPU tests: the oracle pinned against the reference Fortran golden vectors,
brute force and closed forms; the C ABI library loads and exports every
symbol of include/illum_hip.h; host-side helpers."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "cn_golden.npz")


def test_cabi_exports_every_declared_symbol():
    from illum_amd import _lib
    hdr = open(os.path.join(ROOT, "include", "illum_hip.h")).read()
    names = set(re.findall(r"^\s*(?:const\s+)?\w+\s+\*?\s*(illum_\w+)\s*\(", hdr, re.M))
    assert len(names) >= 38
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) == names


def test_oracle_conductionQ_matches_reference_fortran_golden():
    gold = np.load(GOLD)
    for n in gold["names"]:
        z, ti, rhoc, Q = (gold[f"{n}.{k}"] for k in ("z", "ti", "rhoc", "Q"))
        dt, emiss, T0, Fgeo = gold[f"{n}.scalars"]
        nz = len(z)
        T = np.full(nz, T0)
        Ts, Qp = T0, 0.0
        for k, q in enumerate(Q):
            T, Ts, Fs = O.conductionQ(nz, z, dt, Qp, q, T, ti, rhoc, emiss, Ts, Fgeo)
            Qp = q
            # bit-exact: same operations in the same order as conductionQ.f
            assert np.array_equal(T, gold[f"{n}.T"][k]), (n, k)
            assert Ts == gold[f"{n}.Tsurf"][k] and Fs == gold[f"{n}.Fsurf"][k], (n, k)


@pytest.mark.skipif(not O.ref_available(), reason="reference Fortran not built here")
def test_oracle_conductionQ_matches_reference_fortran_live():
    rng = np.random.default_rng(9)
    for _ in range(20):
        nz = int(rng.integers(2, 120))
        z = np.cumsum(rng.uniform(1e-4, 0.05, nz))
        ti = rng.uniform(10, 2000, nz)
        rc = rng.uniform(5e5, 3e6, nz)
        T = rng.uniform(50, 400, nz)
        args = (nz, z, float(rng.uniform(1, 3600)), float(rng.uniform(0, 1400)),
                float(rng.uniform(0, 1400)), T, ti, rc, float(rng.uniform(0.5, 1)),
                float(rng.uniform(50, 400)), float(rng.uniform(0, 1)))
        a = O.conductionQ(*args)
        b = O.ref_conductionQ(*args)
        assert np.array_equal(a[0], b[0]) and a[1] == b[1] and a[2] == b[2]


def test_oracle_thomas_against_dense_solve():
    # The CN step solves A T^{n+1} = r; reconstruct A and r from two calls that
    # differ only in T and check linearity + the dense solution.
    nz = 12
    z = np.cumsum(np.linspace(0.002, 0.03, nz))
    ti, rc, dt, em = 150.0, 1.2e6, 300.0, 0.9
    rng = np.random.default_rng(2)
    T0 = rng.uniform(180, 260, nz)
    Ts = 230.0
    # zero flux + linear diffusion: columns of the update map by unit impulses
    base = O.conductionQ(nz, z, dt, 0.0, 0.0, np.zeros(nz), ti, rc, em, Ts, 0.0)[0]
    M = np.zeros((nz, nz))
    for k in range(nz):
        e = np.zeros(nz)
        e[k] = 1.0
        M[:, k] = O.conductionQ(nz, z, dt, 0.0, 0.0, e, ti, rc, em, Ts, 0.0)[0] - base
    full = O.conductionQ(nz, z, dt, 0.0, 0.0, T0, ti, rc, em, Ts, 0.0)[0]
    np.testing.assert_allclose(full, base + M @ T0, rtol=1e-12, atol=1e-9)
    # the CN update of pure diffusion conserves nothing but is a contraction
    assert np.all(np.abs(np.linalg.eigvals(M)) <= 1 + 1e-12)


def test_oracle_radiative_equilibrium_fixed_point():
    # An isothermal column at eps sigma T^4 = Q stays (nearly) put.
    nz = 30
    z = np.cumsum(np.full(nz, 0.01))
    em, sig = 0.95, 5.6704e-8
    Teq = 250.0
    Q = em * sig * Teq**4
    T = np.full(nz, Teq)
    Ts = Teq
    for _ in range(50):
        T, Ts, Fs = O.conductionQ(nz, z, 600.0, Q, Q, T, 200.0, 1.2e6, em, Ts, 0.0)
    assert abs(Ts - Teq) < 1e-8 and np.max(np.abs(T - Teq)) < 1e-8 and abs(Fs) < 1e-6


def _mesh(subdiv, ncraters, seed=0):
    from illum_amd import api
    return api.synth_ellipsoid(subdiv, ncraters=ncraters, seed=seed)


def test_synth_sizes_and_orientation():
    for s in range(4):
        V, F = _mesh(s, 0)
        assert V.shape == (10 * 4**s + 2, 3) and F.shape == (20 * 4**s, 3)
        g = O.geometry(V, F)
        assert np.all(np.sum(g["normal"] * g["centroid"], 1) > 0)
        # each edge shared by exactly two faces in opposite directions
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
        assert len({tuple(x) for x in e}) == len(e)
        assert {tuple(x) for x in e} == {tuple(x[::-1]) for x in e}
    V1, F1 = _mesh(3, 40, seed=7)
    V2, F2 = _mesh(3, 40, seed=7)
    assert np.array_equal(V1, V2) and np.array_equal(F1, F2)


def test_oracle_bvh_equals_brute_force():
    V, F = _mesh(3, 40, seed=1)
    g = O.geometry(V, F)
    rng = np.random.default_rng(0)
    n = 20000
    o = rng.uniform(-15, 15, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    # grazing rays from facet offsets towards other centroids
    j = rng.integers(0, len(F), 5000)
    i = rng.integers(0, len(F), 5000)
    q = g["centroid"][j] + np.float32(1e-5) * g["normal"][j]
    o = np.concatenate([o, q]).astype(np.float32)
    d = np.concatenate([d, g["centroid"][i] - q]).astype(np.float32)
    h1, t1 = O.closest_hit(g["tri"], o, d, use_bvh=False)
    h2, t2 = O.closest_hit(g["tri"], o, d, use_bvh=True)
    assert np.array_equal(h1, h2)
    assert np.array_equal(t1[h1 >= 0], t2[h1 >= 0])
    b1, _ = O.visibility(g, use_bvh=False)
    b2, _ = O.visibility(g, use_bvh=True)
    assert np.array_equal(b1, b2)


def test_predicate_known_answers():
    # unit right triangle in z = 0, rays straight down
    tri = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32)
    o = np.array([[0.25, 0.25, 1.0], [0.9, 0.9, 1.0], [0.25, 0.25, -1.0], [0.5, 0.5, 2.0]],
                 np.float32)
    d = np.array([[0, 0, -1], [0, 0, -1], [0, 0, -1], [0, 0, -4]], np.float32)
    h, t = O.closest_hit(tri, o, d, use_bvh=False)
    assert list(h) == [0, -1, -1, 0]   # hit, outside, behind origin, on the hypotenuse edge
    assert t[0] == 1.0 and t[3] == 0.5
    # two-sided: from below
    h, t = O.closest_hit(tri, np.array([[0.2, 0.2, -3]], np.float32),
                         np.array([[0, 0, 1]], np.float32), use_bvh=False)
    assert h[0] == 0 and t[0] == 3.0


def test_watertight_shared_edges_never_leak():
    # rays from inside a closed sphere towards every vertex and edge midpoint
    V, F = _mesh(2, 0)
    g = O.geometry(V, F)
    tri = g["tri"].reshape(-1, 3, 3)
    targets = np.concatenate([tri[:, 0], 0.5 * (tri[:, 0] + tri[:, 1]),
                              (tri[:, 0] + tri[:, 1] + tri[:, 2]) / 3]).astype(np.float32)
    o = np.zeros_like(targets) + np.float32([1e-3, -2e-3, 3e-3])
    h, _ = O.closest_hit(g["tri"], o, targets - o, use_bvh=False)
    assert np.all(h >= 0)


def test_form_factor_closed_form_parallel_plates():
    # two small coaxial facing triangles: F = mu mu A_j / (pi d^2) with mu = 1
    d = 3.0
    V = np.array([[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0], [0, 0, d], [0, 0.1, d], [0.1, 0, d]], float)
    F = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    g = O.geometry(V, F)
    assert g["normal"][0][2] == 1 and g["normal"][1][2] == -1
    bits, _ = O.visibility(g, offset=1e-5, use_bvh=False)
    cp, ri, v = O.form_factors(g, bits, albedo=1.0)
    # centroids are coaxial? they are mirror images, so the line is along z
    A = 0.005
    r = np.linalg.norm(g["centroid"][0].astype(float) - g["centroid"][1].astype(float))
    assert list(ri) == [1, 0]
    np.testing.assert_allclose(v, A / (np.pi * r * r), rtol=1e-6)
    # reciprocity A_i F_ij = A_j F_ji
    assert abs(g["area"][1] * v[0] - g["area"][0] * v[1]) < 1e-15


def test_convex_body_has_no_form_factors():
    V, F = _mesh(3, 0)
    g = O.geometry(V, F)
    bits, _ = O.visibility(g)
    cp, ri, v = O.form_factors(g, bits, albedo=1.0)
    assert len(v) == 0


def test_reciprocity_on_cratered_mesh():
    V, F = _mesh(3, 40, seed=3)
    g = O.geometry(V, F)
    bits, _ = O.visibility(g)
    cp, ri, v = O.form_factors(g, bits, albedo=1.0)
    import scipy.sparse as sps
    A = sps.csc_matrix((v, ri, cp), shape=(len(F), len(F))).toarray()
    both = (A > 0) & (A.T > 0)
    Ai = g["area"][:, None]
    lhs, rhs = (Ai * A)[both], (Ai * A).T[both]
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12)
    assert both.sum() > 0
    assert np.all(np.diag(A) == 0)


def test_horizon_flat_plane_and_bisection_count():
    # a flat grid: every horizon lies within theta_eps of pi/2 (just below the
    # plane the own facet / neighbours are hit, just above nothing is)
    n = 8
    xs, ys = np.meshgrid(np.arange(n + 1), np.arange(n + 1))
    V = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)], 1).astype(float)
    F = []
    for a in range(n):
        for b in range(n):
            p = a * (n + 1) + b
            F += [[p, p + 1, p + n + 2], [p, p + n + 2, p + n + 1]]
    F = np.array(F, np.int32)
    g = O.geometry(V, F)
    assert np.all(g["normal"][:, 2] == 1)
    hz, nrays = O.horizons(g, nphi=9, theta_eps=1e-3, offset=1e-5, use_bvh=False)
    assert nrays == len(F) * 9 * 11
    inner = np.abs(g["centroid"][:, 0] - n / 2) < 2
    inner &= np.abs(g["centroid"][:, 1] - n / 2) < 2
    assert np.all(np.abs(hz[:, inner] - np.pi / 2) <= 1e-3 + 1e-12)


def test_direct_unoccluded_closed_form():
    # horizon = pi everywhere: all 99 disk samples count, E = 99/(nphi-1) cos 1372 (AU/d)^2
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    F = np.array([[0, 1, 2]], np.int32)
    g = O.geometry(V, F)
    nphi = 33
    hz = np.full((nphi, 1), np.pi)
    disk = O.fib_spiral(100)
    d = 227390024.0
    sun = np.array([0.0, 0.0, d])
    E = O.direct(g, hz, sun, disk)
    p = g["centroid"][0].astype(float)
    cosz = (sun - p)[2] / np.linalg.norm(sun - p)
    AU = 149.5978707e6
    np.testing.assert_allclose(E[0], 99 / (nphi - 1) * cosz * 1372 * (AU / d) ** 2, rtol=1e-13)
    # horizon 0 with the sun 60 degrees from the zenith: theta > t * dphi for
    # every sample, so nothing counts (a zenith sun would still count through
    # the reference's "+ t * dphi" term, illum.cpp:431-434)
    low = d * np.array([np.sin(np.pi / 3), 0.0, np.cos(np.pi / 3)])
    assert O.direct(g, np.zeros((nphi, 1)), low, disk)[0] == 0.0
    assert O.direct(g, np.full((nphi, 1), np.pi), low, disk)[0] > 0


def test_fib_spiral_closed_form():
    xy = O.fib_spiral(100)
    r = np.hypot(xy[:, 0], xy[:, 1])
    np.testing.assert_allclose(r, np.arange(100) / 101, atol=1e-15)


def test_gauss_seidel_matches_dense_solve():
    V, F = _mesh(3, 40, seed=4)
    g = O.geometry(V, F)
    bits, _ = O.visibility(g)
    cp, ri, v = O.form_factors(g, bits, albedo=0.5)
    import scipy.sparse as sps
    n = len(F)
    A = sps.csc_matrix((v, ri, cp), shape=(n, n)).toarray()
    E = np.random.default_rng(1).uniform(0, 400, n)
    B, k, res = O.gauss_seidel(n, cp, ri, v, E, max_sweeps=100, rel_tol=1e-14)
    Bd = np.linalg.solve(np.eye(n) - A, E)
    np.testing.assert_allclose(B, Bd, rtol=1e-11)
    # one plain sweep pair: GS iterates are the textbook ones
    B1, _, _ = O.gauss_seidel(n, cp, ri, v, E, max_sweeps=1, rel_tol=0)
    L = np.eye(n) - np.tril(A)
    U = -np.triu(A)
    B0 = np.linalg.solve(L, E)
    np.testing.assert_allclose(B1, np.linalg.solve(L, E - U @ B0), rtol=1e-12)


def test_armaio_roundtrip(tmp_path):
    from illum_amd.armaio import load_mat, save_mat
    X = np.random.default_rng(0).normal(size=(7, 3))
    p = tmp_path / "x.bin"
    save_mat(X, p)
    head = open(p, "rb").read(30)
    assert head.startswith(b"ARMA_MAT_BIN_FN008\n7 3\n")
    assert np.array_equal(load_mat(p), X)


def test_oracle_render_sphere_closed_form():
    # Orthographic view of a fine unit icosphere: the camera half-width is 1.1
    # times the centroid radius, so the disk covers pi / (4 * 1.1^2) of a square
    # image up to the polygonal rim; every hit facet faces the camera, and the
    # image is mirror-symmetric about the vertical axis for a view down z.
    from illum_amd import api
    V, F = api.synth_ellipsoid(5, axes=(1.0, 1.0, 1.0), ncraters=0, seed=0)
    g = O.geometry(V, F)
    img = O.render(g, 0, 160, 160, 0.0, 0.0)
    r_cen = np.linalg.norm(g["centroid"].astype(np.float64), axis=1).max()
    frac = np.mean(img >= 0)
    assert abs(frac - np.pi / (4 * 1.1 ** 2)) < 0.01
    assert np.all(g["normal"][img[img >= 0]] @ np.array([0.0, 0.0, 1.0]) > 0)
    assert np.array_equal(img >= 0, (img >= 0)[:, ::-1])
    assert 0.99 < r_cen < 1.0
    # every equal-area pixel looks at the sphere from outside: all hit
    assert np.all(O.render(g, 1, 48, 24, 0.0, 0.0) >= 0)
