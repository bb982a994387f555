"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars: visibility (the union-of-closest-hits sets of compute_F, illum.cpp:42-95),
horizons and ray hits bit-exact; form factors within 1e-12 relative of the
oracle's FP64 values; radiosity at the stated relative residual, recomputed
in FP64 on the host; Crank-Nicolson FP64 within 1e-9 K of the oracle
and the mixed mode within 1e-3 K.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from illum_amd import api as A
    A.default_context()
    return A


def _mesh(api, subdiv, ncraters, seed=0):
    V, F = api.synth_ellipsoid(subdiv, ncraters=ncraters, seed=seed)
    return V, F


def _bits_to_sets(bits, nf):
    return np.unpackbits(bits.view(np.uint8), bitorder="little", axis=1)[:, :nf]


def test_geometry_bitwise(api):
    V, F = _mesh(api, 3, 20)
    m = api.Mesh(api.default_context(), V=V, F=F)
    g = m.geometry()
    o = O.geometry(V, F)
    for k in ("centroid", "normal", "radius", "area", "frame"):
        assert np.array_equal(g[k], o[k]), k
    assert np.array_equal(m.triangles(), o["tri"])


def test_closest_hit_random_rays_bitwise(api):
    V, F = _mesh(api, 3, 30, seed=3)
    ctx = api.IllumContext(V=V, F=F, leaf_size=4)
    o = O.geometry(V, F)
    rng = np.random.default_rng(7)
    n = 50000
    orig = rng.uniform(-14, 14, (n, 3)).astype(np.float32)
    dirs = rng.normal(size=(n, 3)).astype(np.float32)
    # also rays aimed exactly at vertices and edge midpoints (watertightness)
    tri = o["tri"].reshape(-1, 3, 3)
    targets = np.concatenate([tri[:2000, 0], 0.5 * (tri[:2000, 0] + tri[:2000, 1])]).astype(np.float32)
    o2 = np.zeros_like(targets)
    orig = np.concatenate([orig, o2])
    dirs = np.concatenate([dirs, targets])
    hg, tg = ctx.bvh.closest_hit(orig, dirs)
    hb, tb = O.closest_hit(o["tri"], orig, dirs, use_bvh=False)
    assert np.array_equal(hg, hb)
    hit = hb >= 0
    assert np.array_equal(tg[hit], tb[hit])
    # rays from the centre towards vertices/edges of a closed mesh never miss
    assert np.all(hb[n:] >= 0)


@pytest.mark.parametrize("walk", ["default", "packet", "packet4"])
@pytest.mark.parametrize("subdiv,ncraters,albedo", [(3, 40, 0.1), (4, 40, 0.1)])
def test_form_factors_vs_oracle(api, monkeypatch, subdiv, ncraters, albedo, walk):
    if walk != "default":
        monkeypatch.setenv("ILLUM_FF_TRAVERSAL", walk)
    V, F = _mesh(api, subdiv, ncraters, seed=subdiv)
    ctx = api.IllumContext(V=V, F=F, leaf_size=4)
    Fm = ctx.compute_F(albedo, keep_visibility=True)
    o = O.geometry(V, F)
    bits, nrays = O.visibility(o, offset=1e-5)
    gvis = Fm.visibility()[:, : bits.shape[1]]
    assert np.array_equal(gvis, bits), "visibility sets differ"
    assert Fm.rays() == nrays  # same facing-cull survivors
    st = ctx.form_factor_work(1e-5)
    assert st["rays"] == nrays and nrays <= st["pair_tests"] <= len(F) * (len(F) - 1)
    assert st["box_tests"] >= 2 * nrays and st["tri_tests"] >= nrays
    cp, ri, v = O.form_factors(o, bits, albedo=albedo)
    gcp, gri, gv = Fm.csc()
    assert np.array_equal(gcp, cp)
    assert np.array_equal(gri, ri)
    np.testing.assert_allclose(gv, v, rtol=1e-12, atol=0)
    # CSR is the transpose of CSC, rows sorted
    import scipy.sparse as sps
    A = sps.csc_matrix((v, ri, cp), shape=(len(F), len(F))).tocsr()
    A.sort_indices()
    rp, ci, rv = Fm.csr()
    assert np.array_equal(rp, A.indptr)
    assert np.array_equal(ci, A.indices)
    np.testing.assert_allclose(rv, A.data, rtol=1e-12, atol=0)
    assert Fm.nnz > 0


def test_autotuned_bvh_and_traversal_choice_do_not_change_results(api, monkeypatch):
    # Visibility is a property of the triangles, not of the tree: the autotuned
    # leaf size, leaf size 1, the four-wide and two-wide near-first walks, the
    # wavefront packet walk and the stackless threaded walk must all agree.
    V, F = _mesh(api, 4, 40, seed=8)
    ref = api.IllumContext(V=V, F=F, leaf_size=4).compute_F(0.1, keep_visibility=True)
    tuned = api.IllumContext(V=V, F=F, leaf_size=0)
    trials = tuned.bvh.tuning()
    assert [(l, w) for l, w, _ in trials] == [(l, w) for l in (2, 3, 4, 6, 8) for w in (2, 4, 3, 5)]
    assert all(ms > 0 for _, _, ms in trials)
    assert tuned.bvh.info()["leaf_size"] == min(trials, key=lambda x: x[2])[0]
    kv = dict(keep_visibility=True)
    out = [tuned.compute_F(0.1, **kv), api.IllumContext(V=V, F=F, leaf_size=1).compute_F(0.1, **kv)]
    for walk in ("threaded", "wide2", "wide4", "packet", "packet4"):
        monkeypatch.setenv("ILLUM_FF_TRAVERSAL", walk)
        out.append(api.IllumContext(V=V, F=F, leaf_size=4).compute_F(0.1, **kv))
        out.append(api.IllumContext(V=V, F=F, leaf_size=1).compute_F(0.1, **kv))
    r = ref.csc()
    for m in out:
        assert np.array_equal(m.visibility(), ref.visibility())
        c = m.csc()
        assert all(np.array_equal(a, b) for a, b in zip(c, r))


def test_cluster_cull_does_not_change_results(api, monkeypatch):
    # The cluster pass only skips receivers the per-pair test would reject:
    # same visibility and the same rays with it on and off.
    V, F = _mesh(api, 5, 60, seed=11)
    on = api.IllumContext(V=V, F=F).compute_F(0.1, keep_visibility=True)
    monkeypatch.setenv("ILLUM_FF_CLUSTER_CULL", "0")
    off = api.IllumContext(V=V, F=F).compute_F(0.1, keep_visibility=True)
    work_off = api.IllumContext(V=V, F=F).form_factor_work()
    monkeypatch.delenv("ILLUM_FF_CLUSTER_CULL")
    work_on = api.IllumContext(V=V, F=F).form_factor_work()
    assert on.rays() == off.rays() > 0
    assert np.array_equal(on.visibility(), off.visibility())
    nf = len(F)
    assert work_off["pair_tests"] == nf * (nf - 1) and work_off["rays"] == off.rays()
    assert work_on["pair_tests"] < nf * (nf - 1) // 4  # clusters behind the plane
    assert work_on["rays"] == on.rays()
    # Shuffled facets make loose clusters; still bit-exact against the oracle
    # on the shuffled mesh (closest-hit ties go to the smaller index, so the
    # oracle must see the same order).
    V4, F4 = _mesh(api, 4, 40, seed=12)
    F4 = np.ascontiguousarray(F4[np.random.default_rng(3).permutation(len(F4))])
    Fm = api.IllumContext(V=V4, F=F4).compute_F(0.1, keep_visibility=True)
    bits, nrays = O.visibility(O.geometry(V4, F4), offset=1e-5)
    assert np.array_equal(Fm.visibility()[:, : bits.shape[1]], bits)
    assert Fm.rays() == nrays


@pytest.mark.parametrize("subdiv,leaf", [(4, 1), (4, 3), (5, 6)])
def test_per_ray_walk_work_equals_the_oracle_walk_of_the_same_tree(api, monkeypatch, subdiv,
                                                                    leaf):
    # The visibility kernel's roofline prices its work by each ray's own
    # walk: box and triangle tests counted by the instrumented per-lane walk
    # (a bounded walk, then an unbounded one, the threaded walk past 8 stack
    # entries).  The oracle's walk of the same exported tree counts exactly
    # the same and finds the oracle's visibility sets.
    monkeypatch.setenv("ILLUM_FF_TRAVERSAL", "wide2")
    V, F = _mesh(api, subdiv, 40, seed=15)
    ic = api.IllumContext(V=V, F=F, leaf_size=leaf)
    work = ic.form_factor_work(1e-5)
    tree = ic.bvh.export()
    assert tree["roots"].shape == (1, 3) and tree["tris"].shape == (len(F), 12)
    cap = min(ic.bvh.info()["depth"] + 1, 8)
    g = O.geometry(V, F)
    ow, bits = O.walk_work(g, tree, cap=cap, with_bits=True)
    assert ow["rays"] == work["rays"] > 0
    assert ow["box_tests"] == work["box_tests"]
    assert ow["tri_tests"] == work["tri_tests"]
    vis, _ = O.visibility(g)
    assert np.array_equal(bits, vis)


@pytest.mark.parametrize("leaf", [1, 3, 8])
def test_horizon_traversal_choice_does_not_change_results(api, monkeypatch, leaf):
    # Every walk, with the per-lane hit cache on (default) and off, gives the
    # same horizons, at several leaf sizes.
    V, F = _mesh(api, 3, 40, seed=9)
    ctx = api.IllumContext(V=V, F=F, leaf_size=leaf)
    a = ctx.make_horizons(nphi=33).copy()
    for cache in ("1", "0"):
        monkeypatch.setenv("ILLUM_HZ_CACHE", cache)
        for walk in ("default", "threaded", "wide4", "packet"):
            monkeypatch.setenv("ILLUM_HZ_TRAVERSAL", walk)
            b = ctx.make_horizons(nphi=33)
            assert np.array_equal(a, b), (walk, cache)


def test_form_factors_albedo_vector(api):
    V, F = _mesh(api, 3, 40, seed=5)
    ctx = api.IllumContext(V=V, F=F)
    rho = np.random.default_rng(1).uniform(0.05, 0.15, len(F))
    Fm = ctx.compute_F(rho)
    o = O.geometry(V, F)
    bits, _ = O.visibility(o)
    cp, ri, v = O.form_factors(o, bits, albedo_vec=rho)
    gcp, gri, gv = Fm.csc()
    assert np.array_equal(gri, ri)
    np.testing.assert_allclose(gv, v, rtol=1e-12, atol=0)


def test_csr_rank_sort_and_bitmap_paths_agree(api, monkeypatch):
    # CSR(F) is built by scattering the CSC entries into rows, then ordering
    # each row by rank in LDS (rows of up to 1,024 entries) or, forced here
    # for every row, through an LDS bitmap over the body with the values
    # evaluated again; both must give the same arrays, and CSR must be CSC
    # transposed (scipy), with per-facet albedo.
    import scipy.sparse as sps
    V, F = _mesh(api, 4, 40, seed=9)
    ctx = api.IllumContext(V=V, F=F)
    rho = np.random.default_rng(3).uniform(0.05, 0.9, len(F))
    Fa = ctx.compute_F(rho)
    monkeypatch.setenv("ILLUM_FF_CSR_SORT_MAX", "0")
    Fb = ctx.compute_F(rho)
    ra, rb = Fa.csr(), Fb.csr()
    for x, y in zip(ra, rb):
        assert np.array_equal(x, y)
    cp, ri, v = Fa.csc()
    A = sps.csc_matrix((v, ri, cp), shape=(len(F), len(F))).tocsr()
    A.sort_indices()
    assert np.array_equal(A.indptr, ra[0]) and np.array_equal(A.indices, ra[1])
    assert np.array_equal(A.data, ra[2])


def test_long_rows_and_columns_inward_sphere(api, monkeypatch):
    # Facets of a sphere wound inward see nearly every other facet: columns
    # and CSR rows of ~1,279 entries, longer than the one-wave rank sort
    # takes, so rows are ordered through the LDS bitmap.  CSC against the
    # oracle, CSR against scipy's transpose, the all-bitmap path and the
    # all-bitmap path in narrow windows.
    import scipy.sparse as sps
    V, F = _mesh(api, 3, 0)
    F = np.ascontiguousarray(F[:, ::-1])
    ctx = api.IllumContext(V=V, F=F)
    Fa = ctx.compute_F(0.3, keep_visibility=True)
    cp, ri, v = Fa.csc()
    assert np.diff(cp).max() > 1024
    o = O.geometry(V, F)
    bits, _ = O.visibility(o)
    assert np.array_equal(Fa.visibility()[:, : bits.shape[1]], bits)
    ocp, ori, ov = O.form_factors(o, bits, albedo=0.3)
    assert np.array_equal(cp, ocp) and np.array_equal(ri, ori)
    np.testing.assert_allclose(v, ov, rtol=1e-12, atol=0)
    ra = Fa.csr()
    assert np.diff(ra[0]).max() > 1024
    A = sps.csc_matrix((v, ri, cp), shape=(len(F), len(F))).tocsr()
    A.sort_indices()
    assert np.array_equal(A.indptr, ra[0]) and np.array_equal(A.indices, ra[1])
    assert np.array_equal(A.data, ra[2])
    monkeypatch.setenv("ILLUM_FF_CSR_SORT_MAX", "0")
    rb = ctx.compute_F(0.3).csr()
    for x, y in zip(ra, rb):
        assert np.array_equal(x, y)
    # bitmap windows of 3 words (96 sources): every row over many windows
    monkeypatch.setenv("ILLUM_FF_CSR_WINDOW", "3")
    rc = ctx.compute_F(0.3).csr()
    for x, y in zip(ra, rc):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("walk", ["default", "packet"])
@pytest.mark.parametrize("nphi,theta_eps", [(65, 1e-3), (17, 1e-2)])
def test_horizons_bitwise(api, monkeypatch, nphi, theta_eps, walk):
    if walk != "default":
        monkeypatch.setenv("ILLUM_HZ_TRAVERSAL", walk)
    V, F = _mesh(api, 3, 40, seed=2)
    ctx = api.IllumContext(V=V, F=F)
    hz = ctx.make_horizons(nphi=nphi, theta_eps=theta_eps, offset=1e-5)
    o = O.geometry(V, F)
    ho, _ = O.horizons(o, nphi=nphi, theta_eps=theta_eps, offset=1e-5)
    assert hz.shape == (nphi, len(F))
    assert np.array_equal(hz, ho)


def test_direct_radiosity(api):
    V, F = _mesh(api, 3, 40, seed=4)
    ctx = api.IllumContext(V=V, F=F)
    hz = ctx.make_horizons(nphi=65)
    disk = api.fib_spiral(100)
    assert np.array_equal(disk, O.fib_spiral(100))
    rng = np.random.default_rng(3)
    S = 6
    sun = rng.normal(size=(3, S))
    sun = sun / np.linalg.norm(sun, axis=0) * 227390024.0
    E = ctx.get_direct_radiosity(sun, disk)
    Eo = O.direct(O.geometry(V, F), hz, sun, disk)
    np.testing.assert_allclose(E, Eo, rtol=1e-12, atol=1e-12)
    assert E.max() > 0


def test_direct_whole_disk_decision_is_exact(api, monkeypatch):
    # The whole-disk decision only skips samples whose outcome the bound fixes:
    # bit-identical to evaluating every sample, suns at all elevations.
    V, F = _mesh(api, 4, 60, seed=14)
    ctx = api.IllumContext(V=V, F=F)
    ctx.make_horizons(nphi=91)
    disk = api.fib_spiral(100)
    rng = np.random.default_rng(8)
    sun = rng.normal(size=(3, 48))
    sun = sun / np.linalg.norm(sun, axis=0) * 227390024.0
    E = ctx.get_direct_radiosity(sun, disk)
    # the disk samples split over 1..16 lanes per (facet, sun) pair, (0) the
    # open pairs of a wavefront sampled by all its lanes, or (64) the open
    # pairs queued and each sampled by a whole wavefront: an integer count,
    # so the same E whatever the split (one sun and all 48 suns)
    for lanes in ("-16", "-8", "-4", "0", "1", "2", "4", "8", "16", "64"):
        monkeypatch.setenv("ILLUM_DIRECT_LANES", lanes)
        assert np.array_equal(ctx.get_direct_radiosity(sun, disk), E), lanes
        assert np.array_equal(ctx.get_direct_radiosity(sun[:, 7:8], disk)[:, 0], E[:, 7]), lanes
    # every pair open: 245,760 of them overflow the two-pass queue (40,960
    # entries), whose overflow is sampled by each pair's own lane
    monkeypatch.setenv("ILLUM_DIRECT_EXACT", "1")
    for lanes in ("-4", "0", "4", "64"):
        monkeypatch.setenv("ILLUM_DIRECT_LANES", lanes)
        assert np.array_equal(ctx.get_direct_radiosity(sun, disk), E), lanes
    monkeypatch.delenv("ILLUM_DIRECT_LANES")
    Ex = ctx.get_direct_radiosity(sun, disk)
    assert np.array_equal(E, Ex)
    assert (E > 0).mean() > 0.2 and (E == 0).mean() > 0.2


@pytest.mark.parametrize("S", [1, 5, 64])
@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_radiosity_vs_gauss_seidel(api, S, method):
    V, F = _mesh(api, 4, 40, seed=6)
    ctx = api.IllumContext(V=V, F=F)
    Fm = ctx.compute_F(0.5)  # large albedo: a harder solve than the configs
    rng = np.random.default_rng(S)
    E = rng.uniform(0, 500, (len(F), S))
    B, res, iters = api.radiosity_solve(Fm, E if S > 1 else E[:, 0], method=method, rel_tol=1e-10,
                                        max_iter=400)
    B = B.reshape(len(F), -1)
    cp, ri, v = Fm.csc()
    for s in range(min(S, 3)):
        Bo, sweeps, _ = O.gauss_seidel(len(F), cp, ri, v, E[:, s], max_sweeps=200, rel_tol=1e-13)
        np.testing.assert_allclose(B[:, s], Bo, rtol=1e-8, atol=1e-8 * np.abs(Bo).max())
    # FP64 residual of the returned B against the FP64 operator
    import scipy.sparse as sps
    A = sps.csc_matrix((v, ri, cp), shape=(len(F), len(F)))
    r = E - (B - A @ B)
    rel = np.linalg.norm(r, axis=0) / np.linalg.norm(E, axis=0)
    assert rel.max() <= 1.2e-10
    assert res[-1] <= 1e-10


@pytest.mark.parametrize("tol", [1e-10, 1e-8])
@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_radiosity_wide_products_chunking_is_exact(api, monkeypatch, method, tol):
    # S >= 32 runs the wide product over row blocks: every element sums its
    # row in CSR order whatever the chunk width or the blocks' dealing to the
    # XCDs, so forcing narrower chunks (a small cache budget) or another
    # dealing changes nothing; odd S takes one value per lane, S < 64 leaves
    # lanes idle, and mixed solves below S = 128 add each correction to B
    # (B mode).  At 1e-8 the mixed method keeps FP32 remainders (S >= 128),
    # at 1e-10 FP64 residuals.
    V, F = _mesh(api, 4, 40, seed=6)
    ctx = api.IllumContext(V=V, F=F)
    Fm = ctx.compute_F(0.5)
    nf = len(F)
    cp, ri, v = Fm.csc()
    for S in (32, 45, 96, 128, 130, 131, 512):
        E = np.random.default_rng(S).uniform(0, 500, (nf, S))
        monkeypatch.delenv("ILLUM_SPMM_CHUNK_BYTES", raising=False)
        B, res, it = api.radiosity_solve(Fm, E, method=method, rel_tol=tol, max_iter=400)
        assert res[-1] <= tol
        for budget in (nf * 64 * 8 * 1.5, nf * 64 * 8 * 0.5):  # per-lane widths 1..4
            monkeypatch.setenv("ILLUM_SPMM_CHUNK_BYTES", str(budget))
            Bc, resc, itc = api.radiosity_solve(Fm, E, method=method, rel_tol=tol, max_iter=400)
            assert itc == it and np.array_equal(Bc, B) and np.array_equal(resc, res), (S, budget)
        monkeypatch.delenv("ILLUM_SPMM_CHUNK_BYTES", raising=False)
        for xc in ("1", "16"):
            monkeypatch.setenv("ILLUM_RAD_XCD_CHUNK", xc)
            Bc, resc, itc = api.radiosity_solve(Fm, E, method=method, rel_tol=tol, max_iter=400)
            assert itc == it and np.array_equal(Bc, B) and np.array_equal(resc, res), (S, xc)
        monkeypatch.delenv("ILLUM_RAD_XCD_CHUNK", raising=False)
        for s in (0, S - 1):
            Bo, _, _ = O.gauss_seidel(nf, cp, ri, v, E[:, s], max_sweeps=200, rel_tol=1e-13)
            np.testing.assert_allclose(B[:, s], Bo, rtol=10 * tol, atol=10 * tol * np.abs(Bo).max())


def _true_rel_resid(Fm, E, B):
    import scipy.sparse as sps
    cp, ri, v = Fm.csc()
    A = sps.csc_matrix((v, ri, cp), shape=(len(E), len(E)))
    nE = np.linalg.norm(E, axis=0)
    nr = np.linalg.norm(E - (B - A @ B), axis=0)
    return nr, nE


@pytest.mark.parametrize("tol", [1e-8, 1e-9])
def test_radiosity_fp32_remainders_bound_the_true_residual(api, monkeypatch, tol):
    # Ring mode (S >= 128) at tolerances down to 1e-9 keeps the residual as
    # FP32 remainders of the FP16 corrections.  The reported residual adds a
    # bound on the remainders' roundings, so per column it bounds the residual
    # recomputed in FP64 on the host; the solve takes the products the
    # FP64-residual solve takes (ILLUM_RAD_REMAINDER=0) and agrees with it.
    V, F = _mesh(api, 4, 40, seed=6)
    ctx = api.IllumContext(V=V, F=F)
    Fm = ctx.compute_F(0.5)
    nf, S = len(F), 256
    E = np.random.default_rng(3).uniform(0, 500, (nf, S)) * 10.0 ** np.linspace(-20, 20, S)[None, :]
    E[:, 3] = 0.0
    monkeypatch.delenv("ILLUM_RAD_REMAINDER", raising=False)
    B, res, it = api.radiosity_solve(Fm, E, method="mixed", rel_tol=tol, max_iter=400)
    monkeypatch.setenv("ILLUM_RAD_REMAINDER", "0")
    B0, res0, it0 = api.radiosity_solve(Fm, E, method="mixed", rel_tol=tol, max_iter=400)
    assert res[-1] <= tol and res0[-1] <= tol
    assert abs(it - it0) <= 1, (it, it0)
    nr, nE = _true_rel_resid(Fm, E, B)
    ok = nE > 0
    assert np.all(nr[ok] <= res[-1] * nE[ok]), (nr[ok] / nE[ok]).max()
    assert np.all(B[:, 3] == 0.0)
    scale = np.abs(B0).max(axis=0)
    assert np.all(np.abs(B - B0)[:, ok].max(axis=0) <= 4 * tol * scale[ok])


def test_radiosity_fp32_remainder_floor_solves_again_with_fp64_residuals(api, monkeypatch):
    # Forced FP32 remainders with albedo 0.9.  At 1e-12 their rounding bound
    # (about 2^-35 of the first residual) stays below the tolerance: the solve
    # converges with the bound included, which bounds the residual recomputed
    # on the host.  At 1e-14 the bound passes the tolerance after the first
    # product, so the solve is taken again with FP64 residuals and returns
    # exactly what that mode returns (without the second solve it could not
    # converge).
    V, F = _mesh(api, 4, 40, seed=6)
    ctx = api.IllumContext(V=V, F=F)
    Fm = ctx.compute_F(0.9)
    nf, S = len(F), 128
    E = np.random.default_rng(5).uniform(0, 500, (nf, S))
    monkeypatch.setenv("ILLUM_RAD_REMAINDER", "1")
    B, res, it = api.radiosity_solve(Fm, E, method="mixed", rel_tol=1e-12, max_iter=400)
    assert res[-1] <= 1e-12
    nr, nE = _true_rel_resid(Fm, E, B)
    assert np.all(nr <= res[-1] * nE), (nr / nE).max()
    B1, res1, it1 = api.radiosity_solve(Fm, E, method="mixed", rel_tol=1e-14, max_iter=400)
    monkeypatch.setenv("ILLUM_RAD_REMAINDER", "0")
    B0, res0, it0 = api.radiosity_solve(Fm, E, method="mixed", rel_tol=1e-14, max_iter=400)
    assert it1 == it0 and np.array_equal(B1, B0) and np.array_equal(res1, res0)
    assert res1[-1] <= 1e-14


@pytest.mark.parametrize("S", [5, 128])
def test_radiosity_mixed_column_scales_and_folds(api, S):
    # The mixed method rounds each residual column to FP16 after a power-of-two
    # scale taken from that column's norm: columns 60 orders of magnitude apart
    # and an all-zero column must all reach the tolerance, with the residual
    # recomputed in FP64 on the host.  Albedo 0.9 at 1e-12 takes more than 7
    # products, so at S = 128 the ring of corrections is folded into B.
    V, F = _mesh(api, 4, 40, seed=6)
    ctx = api.IllumContext(V=V, F=F)
    Fm = ctx.compute_F(0.9)
    nf = len(F)
    rng = np.random.default_rng(S + 1)
    E = rng.uniform(0, 500, (nf, S)) * 10.0 ** np.linspace(-30, 30, S)[None, :]
    E[:, 1] = 0.0
    tol = 1e-12
    B, res, it = api.radiosity_solve(Fm, E, method="mixed", rel_tol=tol, max_iter=400)
    B64, res64, it64 = api.radiosity_solve(Fm, E, method="fp64", rel_tol=tol, max_iter=400)
    assert it > 7 and res[-1] <= tol
    assert abs(it - it64) <= 1, (it, it64)
    import scipy.sparse as sps
    cp, ri, v = Fm.csc()
    A = sps.csc_matrix((v, ri, cp), shape=(nf, nf))
    r = E - (B - A @ B)
    nE = np.linalg.norm(E, axis=0)
    assert np.all(np.linalg.norm(r[:, nE > 0], axis=0) <= 1.2 * tol * nE[nE > 0])
    assert np.all(B[:, 1] == 0.0)
    scale = np.abs(B64).max(axis=0)
    ok = nE > 0
    assert np.all(np.abs(B - B64)[:, ok].max(axis=0) <= 1e-10 * scale[ok])


@pytest.mark.parametrize("S", [1, 128])
@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_radiosity_not_converged_reports_and_writes_B(api, S, method):
    # Stopped by max_iter above the tolerance: status ILLUM_ERR_NOT_CONVERGED,
    # with B written.  Mixed: B is the sum of the corrections so far, whose
    # FP64 residual is exactly the last one reported.  FP64: B is the newest
    # iterate, whose residual is below the last one reported (that of the
    # iterate before it).
    import scipy.sparse as sps
    C, L = api.C, api._lib
    V, F = _mesh(api, 3, 20, seed=4)
    ctx = api.IllumContext(V=V, F=F)
    Fm = ctx.compute_F(0.9)
    nf = len(F)
    E = np.random.default_rng(S).uniform(0, 500, (nf, S))
    Ecm = np.ascontiguousarray(E.T)
    Bcm = np.zeros_like(Ecm)
    resid = np.zeros(3)
    iters = C.c_int()
    m = {"fp64": L.ILLUM_RAD_JACOBI_FP64, "mixed": L.ILLUM_RAD_MIXED}[method]
    with pytest.raises(L.IllumError) as ei:
        L.call("illum_radiosity_solve", Fm.h, api._ptr(Ecm), S, m, 1e-15, 2, api._ptr(Bcm),
               api._ptr(resid), C.byref(iters), 0)
    assert ei.value.code == L.ILLUM_ERR_NOT_CONVERGED
    assert iters.value == 2 and resid[1] > 1e-15 and resid[1] < resid[0]
    cp, ri, v = Fm.csc()
    A = sps.csc_matrix((v, ri, cp), shape=(nf, nf))
    B = Bcm.T
    rel = (np.linalg.norm(E - (B - A @ B), axis=0) / np.linalg.norm(E, axis=0)).max()
    if method == "mixed":
        assert abs(rel - resid[1]) <= 1e-9 * resid[1]
    else:
        assert rel < resid[1]


@pytest.mark.parametrize("S", [128, 200])
@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_radiosity_on_a_given_csr_matrix_vs_dense_solve(api, S, method):
    # F from illum_sparse_create_csr (not a form-factor assembly): n = 1,001
    # rows, so the wide product's last group of rows is partial; rows of
    # random nonnegative values summing to 0.5.  B against a dense solve.
    rng = np.random.default_rng(S)
    n = 1001
    rowptr, colidx, val = [0], [], []
    for i in range(n):
        k = int(rng.integers(0, 40))
        cols = np.sort(rng.choice(np.delete(np.arange(n), i), size=k, replace=False))
        w = rng.uniform(0, 1, k)
        colidx.extend(cols.tolist())
        val.extend((0.5 * w / max(w.sum(), 1e-300)).tolist())
        rowptr.append(len(colidx))
    Fm = api.SparseMatrix.from_csr(api.default_context(), n, rowptr, colidx, val)
    E = rng.uniform(0, 500, (n, S))
    B, res, it = api.radiosity_solve(Fm, E, method=method, rel_tol=1e-12, max_iter=400)
    A = np.zeros((n, n))
    for i in range(n):
        A[i, colidx[rowptr[i]:rowptr[i + 1]]] = val[rowptr[i]:rowptr[i + 1]]
    Bd = np.linalg.solve(np.eye(n) - A, E)
    assert res[-1] <= 1e-12
    np.testing.assert_allclose(B, Bd, rtol=1e-10, atol=1e-10 * np.abs(Bd).max())


def _cn_case(nf, nz, seed):
    rng = np.random.default_rng(seed)
    z = 0.5 * (np.exp(np.linspace(0, 3, nz)) - 1 + 0.01) / (np.exp(3) - 1 + 0.01)
    Q = rng.uniform(0, 600, (64, nf))
    return z, Q


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("mixed", 1e-3)])
def test_thermal_vs_oracle(api, precision, tol):
    nf, nz = 700, 40
    z, Q = _cn_case(nf, nz, 11)
    dt = 5.342 * 3600 / 256
    th = api.ThermalModel(nf, z, 200.0, 1.19177e6, 233.0, emissivity=0.95, precision=precision)
    th.run(dt, Q[:32])
    for k in range(32, 64):
        th.step(dt, Q[k])
    T0 = np.full((nz + 1, nf), 233.0)
    To, _, Fo, _ = O.thermal_run([0, nz], z, np.full(nz, 200.0), np.full(nz, 1.19177e6), T0,
                                 np.zeros(nf), Q, dt, 0.95)
    T = th.T
    Ts, Fs = th.surface()
    dT, dF = np.max(np.abs(T - To)), np.max(np.abs(th.get_radiosity() + Fo))
    print(f"{precision}: max |T - oracle| {dT:.3e} K, max |Fsurf - oracle| {dF:.3e} W/m^2")
    assert dT <= tol
    assert np.max(np.abs(Ts - To[0])) <= tol
    # Fsurf = -k (T1 - Tsurf) / z1 amplifies temperature errors by k / z1
    # (about 130 W m^-2 K^-1 here), so its bound follows the temperature's
    assert dF <= 200 * tol if precision == "mixed" else np.allclose(
        th.get_radiosity(), -Fo, rtol=1e-9, atol=1e-6)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("mixed", 1e-3)])
def test_thermal_ragged_layers_and_ti_scale(api, precision, tol):
    """Grids of 32..150 layers cover every engine: register kernels for 32, 64
    and 96 layers, the 128-layer kernel (FP64 multipliers in scratch) and the
    streaming kernel beyond 128 layers."""
    rng = np.random.default_rng(5)
    nf = 700
    depths = (32, 57, 96, 120, 150)
    grids = []
    for nz in depths:
        grids.append(0.5 * (np.exp(np.linspace(0, 3, nz)) - 1 + 0.01) / (np.exp(3) - 1 + 0.01))
    gid = rng.integers(0, len(depths), nf).astype(np.int32)
    s = np.exp(rng.normal(0, 0.5, nf))
    dt = 60.0
    Q = rng.uniform(0, 600, (40, nf))
    th = api.ThermalModel(nf, grids, 200.0, 1.19177e6, 250.0, emissivity=0.95,
                          grid_of_facet=gid, ti_scale=s, precision=precision)
    th.run(dt, Q)
    goff = np.cumsum([0] + [len(g) for g in grids])
    zc = np.concatenate(grids)
    nzmax = max(depths)
    T0 = np.full((nzmax + 1, nf), 250.0)
    To, _, Fo, _ = O.thermal_run(goff, zc, np.full(len(zc), 200.0), np.full(len(zc), 1.19177e6),
                                 T0, np.zeros(nf), Q, dt, 0.95, grid=gid, ti_scale=s)
    T = th.T
    mask = ~np.isnan(T)
    nz = np.array([len(grids[g]) for g in gid])
    assert np.array_equal(mask, np.arange(nzmax + 1)[:, None] <= nz[None, :])
    err = np.abs(T[mask] - To[mask])
    print(f"{precision}: max |T - oracle| = {err.max():.3e} K")
    assert err.max() <= tol
    # Fsurf = -k (T1 - Tsurf) / z1 amplifies temperature errors by k / z1
    # (about 130 W m^-2 K^-1 on these grids)
    Fs = th.surface()[1]
    assert np.max(np.abs(Fs - Fo)) <= tol * 200


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_thermal_run_with_held_flux_equals_single_steps(api, precision):
    """ILLUM_Q_HELD: one flux row held for n steps in one launch sequence gives
    the state n single steps with that flux give, bit for bit, over every
    engine (ragged grids of 32..150 layers), host and device flux."""
    rng = np.random.default_rng(11)
    nf = 700
    depths = (32, 57, 96, 120, 150)
    grids = [0.5 * (np.exp(np.linspace(0, 3, nz)) - 1 + 0.01) / (np.exp(3) - 1 + 0.01)
             for nz in depths]
    gid = rng.integers(0, len(depths), nf).astype(np.int32)
    s = np.exp(rng.normal(0, 0.5, nf))
    kw = dict(emissivity=0.95, grid_of_facet=gid, ti_scale=s, precision=precision)
    a = api.ThermalModel(nf, grids, 200.0, 1.19177e6, 250.0, **kw)
    b = api.ThermalModel(nf, grids, 200.0, 1.19177e6, 250.0, **kw)
    Qs = [rng.uniform(0, 600, nf) for _ in range(3)]
    counts = (5, 1, 12)
    for Q, n in zip(Qs, counts):
        for _ in range(n):
            a.step(60.0, Q)
    for i, (Q, n) in enumerate(zip(Qs, counts)):
        if i == 2:
            Qd = api.DeviceArray.from_host(api.default_context(), Q)
            b.run_held(60.0, Qd, n)
        else:
            b.run_held(60.0, Q, n)
    Ta, Tb = a.T, b.T
    mask = ~np.isnan(Ta)
    assert np.array_equal(np.isnan(Tb), ~mask)
    assert np.array_equal(Ta[mask], Tb[mask])
    assert all(np.array_equal(x, y) for x, y in zip(a.surface(), b.surface()))


def test_thermal_get_radiosity_accumulates(api):
    rng = np.random.default_rng(9)
    nf = 300
    z = 0.5 * (np.exp(np.linspace(0, 3, 40)) - 1 + 0.01) / (np.exp(3) - 1 + 0.01)
    th = api.ThermalModel(nf, z, 200.0, 1.19177e6, 230.0, emissivity=0.95)
    th.step(60.0, rng.uniform(0, 600, nf))
    Fs = th.surface()[1]
    assert np.array_equal(th.get_radiosity(), -Fs)
    E = rng.uniform(0, 10, nf)
    out = E.copy()
    th.get_radiosity(out=out, accumulate=True)
    assert np.array_equal(out, E - Fs)
    ctx = api.default_context()
    Ed = api.DeviceArray.from_host(ctx, E)
    th.get_radiosity(out=Ed, accumulate=True)
    assert np.array_equal(Ed.to_host(), E - Fs)
    th.get_radiosity(out=Ed)
    assert np.array_equal(Ed.to_host(), -Fs)


@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_async_time_loop_matches_synchronous(api, method):
    """The coupled time step (direct -> E -= F_surf -> solve -> CN step,
    illum_cli.cpp:403-456) on device buffers: ILLUM_ASYNC calls, which queue
    without waiting, give bit-identical temperatures and product counts to
    the same calls made synchronously; the solve's B matches the host-buffer
    solve of the same E."""
    from illum_amd import _lib
    V, F = _mesh(api, 3, 20, seed=21)
    ctx = api.IllumContext(V=V, F=F)
    Fm = ctx.compute_F(0.12)
    hz = ctx.make_horizons(nphi=91)
    nf = len(F)
    c = api.default_context()
    hzd = api.DeviceArray.from_host(c, np.ascontiguousarray(np.asarray(hz).T))
    disk = api.DeviceArray.from_host(c, np.ascontiguousarray(api.fib_spiral(100).T))
    z = 0.5 * (np.exp(np.linspace(0, 3, 30)) - 1 + 0.01) / (np.exp(3) - 1 + 0.01)
    ang = 2 * np.pi * np.arange(6) / 16
    suns = [api.DeviceArray.from_host(c, 2.27e8 * np.array([np.cos(a) * 0.94, np.sin(a) * 0.94, 0.34]))
            for a in ang]
    m = {"fp64": _lib.ILLUM_RAD_JACOBI_FP64, "mixed": _lib.ILLUM_RAD_MIXED}[method]
    runs = {}
    for asyn in (0, _lib.ILLUM_ASYNC):
        dev = _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_OUT_ON_DEVICE | asyn
        th = api.ThermalModel(nf, z, 200.0, 1.19177e6, 200.0, emissivity=0.95,
                              precision=method)
        E = api.DeviceArray(c, (nf,))
        B = api.DeviceArray(c, (nf,))
        counts, Es = [], []
        for k in range(len(suns)):
            _lib.call("illum_direct_radiosity", ctx.mesh.h, api._ptr(hzd), 91, api._ptr(suns[k]),
                      1, api._ptr(disk), 100, api._ptr(E), dev)
            th.get_radiosity(out=E, accumulate=True)
            Es.append(E.to_host())
            resid = np.zeros(101)
            it = api.C.c_int()
            _lib.call("illum_radiosity_solve", Fm.h, api._ptr(E), 1, m, 1e-8, 100, api._ptr(B),
                      api._ptr(resid), api.C.byref(it), dev)
            counts.append(it.value)
            _lib.call("illum_thermal_step", th.h, 3600.0, api._ptr(B), _lib.ILLUM_IN_ON_DEVICE | asyn)
        c.synchronize()
        runs[asyn] = (th.T.copy(), counts, Es[-1], B.to_host())
        th.close()
    (T0, n0, E0, B0), (T1, n1, E1, B1) = runs[0], runs[_lib.ILLUM_ASYNC]
    assert n0 == n1 and min(n0) >= 1
    assert np.array_equal(T0, T1, equal_nan=True)
    assert np.array_equal(B0, B1)
    Bh, res, _ = api.radiosity_solve(Fm, E1, method=method, rel_tol=1e-8, max_iter=100)
    assert np.array_equal(Bh, B1)
    assert res[-1] <= 1e-8


@pytest.mark.parametrize("method,tol_K", [("fp64", 1e-5), ("mixed", 1e-3)])
def test_c1_coupled_loop_5120_facets_256_steps_vs_oracle(api, method, tol_K):
    """The reference's CPU-runnable configuration C1 (illum_amd/configs.py):
    5,120 facets (subdivision 4, 40 craters), albedo 0.1, 64 azimuth
    intervals (nphi = 65), one sun, 256 coupled steps on 40 layers
    (illum_cli.cpp:403-456: direct -> E -= F_surf -> solve -> CN step) with
    the solve to 1e-10, on device buffers, against the oracle's replay with
    Gauss-Seidel to 1e-10.  Horizons bit-exact; temperatures within tol_K of
    the oracle."""
    from illum_amd import _lib, configs
    V, F = configs.mesh("C1")
    assert len(F) == 5120
    ctx = api.IllumContext(V=V, F=F)
    Fm = ctx.compute_F(0.1)
    nphi = 65
    hz = ctx.make_horizons(nphi=nphi)
    g = O.geometry(V, F)
    ho, _ = O.horizons(g, nphi=nphi)
    assert np.array_equal(hz, ho)
    nf = len(F)
    nz, nsteps = 40, 256
    z = configs.layer_grid(nz)
    dt = configs.ROT_PERIOD_S / nsteps
    sun = np.stack([configs.orbit_sun(k, nsteps) for k in range(nsteps)])
    c = api.default_context()
    hzd = api.DeviceArray.from_host(c, np.ascontiguousarray(np.asarray(hz).T))
    disk_h = api.fib_spiral(100)
    disk = api.DeviceArray.from_host(c, np.ascontiguousarray(disk_h.T))
    m = {"fp64": _lib.ILLUM_RAD_JACOBI_FP64, "mixed": _lib.ILLUM_RAD_MIXED}[method]
    dev = _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_OUT_ON_DEVICE | _lib.ILLUM_ASYNC
    th = api.ThermalModel(nf, z, configs.TI, configs.RHOC, 200.0, emissivity=configs.EMISS,
                          precision=method)
    E = api.DeviceArray(c, (nf,))
    B = api.DeviceArray(c, (nf,))
    suns = [api.DeviceArray.from_host(c, sun[k]) for k in range(nsteps)]
    resid = np.zeros(101)
    it = api.C.c_int()
    for k in range(nsteps):
        _lib.call("illum_direct_radiosity", ctx.mesh.h, api._ptr(hzd), nphi, api._ptr(suns[k]), 1,
                  api._ptr(disk), 100, api._ptr(E), dev)
        th.get_radiosity(out=E, accumulate=True)
        _lib.call("illum_radiosity_solve", Fm.h, api._ptr(E), 1, m, 1e-10, 100, api._ptr(B),
                  api._ptr(resid), api.C.byref(it), dev)
        _lib.call("illum_thermal_step", th.h, dt, api._ptr(B),
                  _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_ASYNC)
    c.synchronize()
    Tg = th.T.copy()
    th.close()
    # oracle replay
    bits, _ = O.visibility(g)
    cp, ri, v = O.form_factors(g, bits, albedo=0.1)
    T = np.full((nz + 1, nf), 200.0)
    Qprev = np.zeros(nf)
    Fs = np.zeros(nf)
    for k in range(nsteps):
        Eo = -Fs + O.direct(g, ho, sun[k], disk_h)
        Bo, _, _ = O.gauss_seidel(nf, cp, ri, v, Eo, max_sweeps=200, rel_tol=1e-10)
        T, Qprev, Fs, _ = O.thermal_run([0, nz], z, np.full(nz, configs.TI),
                                        np.full(nz, configs.RHOC), T, Qprev, Bo[None, :], dt,
                                        configs.EMISS)
    assert np.isfinite(Tg).all()
    err = np.abs(Tg - T).max()
    print(f"C1 {method}: max |T - oracle| {err:.2e} K over 256 coupled steps")
    assert err <= tol_K


def test_conductionQ_dropin_vs_reference_fortran_golden(api):
    import os
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "cn_golden.npz"))
    for n in gold["names"]:
        z, ti, rhoc, Q = (gold[f"{n}.{k}"] for k in ("z", "ti", "rhoc", "Q"))
        dt, emiss, T0, Fgeo = gold[f"{n}.scalars"]
        nz = len(z)
        T = np.full(nz, T0)
        Ts = T0
        Qp = 0.0
        for k, q in enumerate(Q[:8]):
            T, Ts, Fs = api.conductionQ(nz, z, dt, Qp, q, T, ti, rhoc, emiss, Ts, Fgeo)
            Qp = q
            np.testing.assert_allclose(T, gold[f"{n}.T"][k], rtol=0, atol=1e-10)
            assert abs(Ts - gold[f"{n}.Tsurf"][k]) <= 1e-10
            assert abs(Fs - gold[f"{n}.Fsurf"][k]) <= 1e-8 * max(1.0, abs(Fs))


def test_multibody_batch_vs_oracle_per_body(api, monkeypatch):
    """Bodies of uneven size in one launch sequence (the multi-body batch
    configuration, scaled down): each body's visibility, form factors,
    horizons and scattered radiosity equal the oracle run on that body alone,
    and nothing couples bodies (F is block-diagonal)."""
    import scipy.sparse as sps
    rng = np.random.default_rng(21)
    bodies = []
    for k, sub in enumerate((2, 3, 4, 3, 2)):
        ax = 13.0 * rng.uniform(0.6, 1.0, 3)
        bodies.append(api.synth_ellipsoid(sub, axes=tuple(ax), ncraters=20, seed=100 + k))
    V, F, off = api.concat_bodies(bodies)
    nf = len(F)
    ctx = api.IllumContext(V=V, F=F, bodies=off, leaf_size=0)  # autotuned forest
    Fm = ctx.compute_F(0.1, keep_visibility=True)
    vis = _bits_to_sets(Fm.visibility(), nf)
    cp, ri, v = Fm.csc()
    A = sps.csc_matrix((v, ri, cp), shape=(nf, nf))
    hz = ctx.make_horizons(nphi=17).copy()
    # the wavefront walk: lanes of one wavefront span body boundaries here
    monkeypatch.setenv("ILLUM_HZ_TRAVERSAL", "packet")
    assert np.array_equal(ctx.make_horizons(nphi=17), hz)
    monkeypatch.delenv("ILLUM_HZ_TRAVERSAL")
    E = rng.uniform(0, 500, (nf, 3))
    B, res, _ = api.radiosity_solve(Fm, E, method="mixed", rel_tol=1e-10, max_iter=400)
    rays = 0
    for b, (Vb, Fb) in enumerate(bodies):
        lo, hi = int(off[b]), int(off[b + 1])
        g = O.geometry(Vb, Fb)
        bits, nr = O.visibility(g, offset=1e-5)
        rays += nr
        assert np.array_equal(vis[lo:hi, lo:hi], _bits_to_sets(bits, hi - lo)), f"body {b}"
        assert not vis[lo:hi, :lo].any() and not vis[lo:hi, hi:].any()
        ocp, ori, ov = O.form_factors(g, bits, albedo=0.1)
        blk = A[:, lo:hi].tocsc()
        blk.sort_indices()
        assert np.array_equal(blk.indptr, ocp) and np.array_equal(blk.indices - lo, ori)
        np.testing.assert_allclose(blk.data, ov, rtol=1e-12, atol=0)
        ho, _ = O.horizons(g, nphi=17, theta_eps=1e-3, offset=1e-5)
        assert np.array_equal(hz[:, lo:hi], ho), f"body {b} horizons"
        Bo, _, _ = O.gauss_seidel(hi - lo, ocp, ori, ov, E[lo:hi, 0], max_sweeps=200,
                                  rel_tol=1e-13)
        np.testing.assert_allclose(B[lo:hi, 0], Bo, rtol=1e-8, atol=1e-8 * np.abs(Bo).max())
    assert Fm.rays() == rays
    assert res[-1] <= 1e-10


def test_bvh_built_for_another_partition_is_rejected(api):
    V, F = _mesh(api, 2, 5, seed=1)
    V2, F2, off = api.concat_bodies([(V, F), (V, F)])
    m = api.Mesh(api.default_context(), V=V2, F=F2)
    t = api.BVH(m, 4)
    api._lib.call("illum_mesh_set_bodies", m.h, api._ptr(off), len(off) - 1)
    h = api.C.c_void_p()
    with pytest.raises(api._lib.IllumError):
        api._lib.call("illum_form_factors", t.h, 0.1, None, 1e-5, api.C.byref(h))


@pytest.mark.parametrize("mode,az,el", [(0, 30.0, 20.0), (0, 0.0, 90.0), (1, 0.3, 0.1)])
def test_render_vs_oracle(api, mode, az, el):
    V, F = _mesh(api, 4, 40, seed=13)
    ctx = api.IllumContext(V=V, F=F)
    data = np.random.default_rng(2).uniform(100, 400, len(F))
    name = "orthographic" if mode == 0 else "equal_area_cylindrical"
    fac, val = ctx.render(data, mode=name, width=96, height=64, az=az, el=el, minval=150,
                          maxval=350)
    fo = O.render(O.geometry(V, F), mode, 96, 64, az, el)
    assert np.array_equal(fac, fo)
    hit = fo >= 0
    np.testing.assert_array_equal(val[hit], np.clip(data[fo[hit]], 150, 350))
    assert np.all(np.isnan(val[~hit]))


def test_render_sphere_known_answers(api):
    # A fine icosphere viewed orthographically fills pi / (4 * 1.1^2) of a
    # square image (camera half-width = 1.1 x the centroid radius), and each
    # pixel sees a facet facing the camera; the equal-area projection sees the
    # sphere from all around, every pixel hitting the facet in its direction.
    V, F = api.synth_ellipsoid(5, axes=(1.0, 1.0, 1.0), ncraters=0, seed=0)
    ctx = api.IllumContext(V=V, F=F)
    n = ctx.mesh.geometry()["normal"].astype(np.float64)
    fac, _ = ctx.render(None, width=200, height=200, az=40.0, el=70.0)
    frac = np.mean(fac >= 0)
    assert abs(frac - np.pi / (4 * 1.1 ** 2)) < 0.01
    az, el = np.deg2rad(40.0), np.deg2rad(70.0)
    toward_cam = np.array([np.cos(az) * np.sin(el), np.sin(az) * np.sin(el), np.cos(el)])
    assert np.all(n[fac[fac >= 0]] @ toward_cam > 0)
    fac2, _ = ctx.render(None, mode="equal_area_cylindrical", width=64, height=32)
    assert np.all(fac2 >= 0)
