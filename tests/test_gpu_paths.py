"""GPU tests of the input paths and error paths: the OBJ loader
(obj_util.cpp:8-96), horizons against a separate occluder mesh
(illum.cpp:333-349), the NaN and zero-pivot status codes (illum.cpp:121-123,
tridag.for:10-12), device-resident right-hand-side blocks, and consecutive
solves on one matrix.  Bars as in test_gpu_parity.py: geometry and horizons
bit-exact against the oracle, radiosity at the requested residual."""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OBJ = os.path.join(ROOT, "tests", "golden", "two_shapes.obj")


@pytest.fixture(scope="module")
def api():
    from illum_amd import api as A
    A.default_context()
    return A


def _write_obj(path, V, F, scale=100.0):
    with open(path, "w") as f:
        for v in V * scale:
            f.write("v %r %r %r\n" % tuple(float(x) for x in v))
        for t in F + 1:
            f.write("f %d %d %d\n" % tuple(t))


def test_obj_loader_matches_oracle_get_objects(api):
    ic = api.IllumContext(obj_path=OBJ)
    assert ic.get_num_faces() == 10  # first shape only
    g = ic.mesh.geometry()
    o = O.get_objects(OBJ)
    for k in ("centroid", "normal", "radius", "area", "frame"):
        assert np.array_equal(g[k], o[k]), k
    assert np.array_equal(ic.mesh.triangles(), o["tri"])


def test_obj_synthetic_mesh_form_factors_and_horizons(api, tmp_path):
    # a cratered ellipsoid written as OBJ in hundreds of units: the loader's
    # /100 gives back the same scale (in FP32), and the whole path runs on it
    V, F = api.synth_ellipsoid(3, ncraters=40, seed=17)
    p = tmp_path / "body.obj"
    _write_obj(p, V, F)
    ic = api.IllumContext(obj_path=str(p))
    o = O.get_objects(str(p))
    assert np.array_equal(ic.mesh.geometry()["normal"], o["normal"])
    Fm = ic.compute_F(0.1, keep_visibility=True)
    bits, _ = O.visibility(o)
    assert np.array_equal(Fm.visibility()[:, : bits.shape[1]], bits)
    cp, ri, v = O.form_factors(o, bits, albedo=0.1)
    gcp, gri, gv = Fm.csc()
    assert np.array_equal(gcp, cp) and np.array_equal(gri, ri)
    np.testing.assert_allclose(gv, v, rtol=1e-12, atol=0)
    hz = ic.make_horizons(nphi=33)
    ho, _ = O.horizons(o, nphi=33)
    assert np.array_equal(np.asarray(hz), ho)


def test_horizons_against_an_occluder_mesh(api, tmp_path):
    # facets of one body, rays cast against another mesh: the same ellipsoid
    # with other craters, scaled up by 1 % so it rises around the first
    V, F = api.synth_ellipsoid(3, ncraters=40, seed=2)
    Vo, Fo = api.synth_ellipsoid(4, ncraters=60, seed=5)
    p = tmp_path / "occ.obj"
    _write_obj(p, Vo * 1.01, Fo)
    ic = api.IllumContext(V=V, F=F)
    ic.set_horizon_obj_file(str(p))
    hz = np.asarray(ic.make_horizons(nphi=33))
    g = O.geometry(V, F)
    occ = O.get_objects(str(p))
    ho, _ = O.horizons(g, nphi=33, occ_tri=occ["tri"])
    assert np.array_equal(hz, ho)
    # the occluder matters: the facets' own-mesh map differs
    own, _ = O.horizons(g, nphi=33)
    assert not np.array_equal(ho, own)


def test_nan_form_factor_is_reported(api):
    from illum_amd import _lib
    V, F = api.synth_ellipsoid(3, ncraters=40, seed=5)
    ic = api.IllumContext(V=V, F=F)
    g = O.geometry(V, F)
    bits, _ = O.visibility(g)
    cp, ri, v = O.form_factors(g, bits, albedo=1.0)
    rho = np.full(len(F), 0.1)
    rho[ri[0]] = np.nan  # a receiver some source sees
    with pytest.raises(FloatingPointError):
        O.form_factors(g, bits, albedo_vec=rho)
    with pytest.raises(_lib.IllumError) as e:
        ic.compute_F(rho)
    assert e.value.code == _lib.ILLUM_ERR_NAN
    # the context stays usable
    assert ic.compute_F(0.1).nnz == len(v)


def test_zero_pivot_is_reported(api):
    from illum_amd import _lib
    from tests.test_cpu_pins import zero_pivot_case
    c = zero_pivot_case()
    with pytest.raises(_lib.IllumError) as e:
        api.conductionQ(**c)
    assert e.value.code == _lib.ILLUM_ERR_ZERO_PIVOT
    # one ulp away: the GPU step equals the oracle's
    c2 = dict(c, emiss=float(np.nextafter(c["emiss"], 0)))
    T, Ts, Fs = api.conductionQ(**c2)
    To, Tso, Fso = O.conductionQ(**c2)
    assert np.array_equal(T, To) and Ts == Tso and Fs == Fso
    # the model form: same column as a two-layer facet
    th = api.ThermalModel(1, c["z"], 1.0, 1.0, 1.0, emissivity=c["emiss"])
    with pytest.raises(_lib.IllumError) as e:
        th.step(1.0, np.array([20.0]))
    assert e.value.code == _lib.ILLUM_ERR_ZERO_PIVOT


@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_radiosity_device_block_of_right_hand_sides(api, method):
    V, F = api.synth_ellipsoid(3, ncraters=40, seed=6)
    ic = api.IllumContext(V=V, F=F)
    Fm = ic.compute_F(0.5)
    n, S = len(F), 6
    E = np.random.default_rng(2).uniform(0, 400, (n, S))
    Bh, res_h, it_h = api.radiosity_solve(Fm, E, method=method, rel_tol=1e-10)
    ctx = api.default_context()
    # device layout: (S, N) row-major = the C ABI's N x S column-major
    Ed = api.DeviceArray.from_host(ctx, np.ascontiguousarray(E.T))
    Bd = api.DeviceArray(ctx, (S, n))
    _, res_d, it_d = api.radiosity_solve(Fm, Ed, method=method, rel_tol=1e-10, B=Bd)
    ctx.synchronize()
    assert it_d == it_h and np.array_equal(res_d, res_h)
    assert np.array_equal(Bd.to_host().T, Bh)
    # shapes that do not hold N x S values are refused before the GPU call
    from illum_amd import _lib
    with pytest.raises(_lib.IllumError):
        api.radiosity_solve(Fm, api.DeviceArray(ctx, (n, S)), method=method, B=Bd)
    with pytest.raises(_lib.IllumError):
        api.radiosity_solve(Fm, Ed, method=method, B=api.DeviceArray(ctx, (S, n - 1)))


@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_consecutive_solves_report_their_own_products(api, method):
    # a converged solve followed by one stopped at max_iter = 1 on the same
    # matrix and context: the second reports 1 product and 1 residual, not
    # the first solve's count (the device control block is reset per solve)
    from illum_amd import _lib
    V, F = api.synth_ellipsoid(3, ncraters=40, seed=6)
    ic = api.IllumContext(V=V, F=F)
    Fm = ic.compute_F(0.5)
    E = np.random.default_rng(3).uniform(0, 400, (len(F), 3))
    B, res, it = api.radiosity_solve(Fm, E, method=method, rel_tol=1e-12, max_iter=200)
    assert it > 2 and res[it - 1] <= 1e-12
    with pytest.raises(_lib.IllumError) as e:
        api.radiosity_solve(Fm, E, method=method, rel_tol=1e-12, max_iter=1)
    assert e.value.code == _lib.ILLUM_ERR_NOT_CONVERGED
    B2, res2, it2 = api.radiosity_solve(Fm, E, method=method, rel_tol=1e-12, max_iter=200)
    assert it2 == it and np.array_equal(res2, res) and np.array_equal(B2, B)


def test_autotuned_launch_shapes_do_not_change_results(api, monkeypatch):
    # The autotuner picks the direct kernel's lanes per pair and the wide
    # SpMM's right-hand sides per lane by timing; every choice must give the
    # same E and B as each forced choice.
    from illum_amd import configs
    V, F = api.synth_ellipsoid(4, ncraters=40, seed=23)
    ic = api.IllumContext(V=V, F=F)
    Fm = ic.compute_F(0.3)
    ic.make_horizons(nphi=33)
    disk = api.fib_spiral(100)
    sun = configs.hemisphere_suns(136, 4)
    E = ic.get_direct_radiosity(sun, disk)
    for L in (-16, -8, -4, 0, 1, 2, 4, 8, 16, 64):
        monkeypatch.setenv("ILLUM_DIRECT_LANES", str(L))
        assert np.array_equal(ic.get_direct_radiosity(sun, disk), E), L
    monkeypatch.delenv("ILLUM_DIRECT_LANES")
    for method in ("fp64", "mixed"):
        B, res, it = api.radiosity_solve(Fm, E, method=method, rel_tol=1e-10)
        # a chunk budget of 1 byte forces one right-hand side per lane
        monkeypatch.setenv("ILLUM_SPMM_CHUNK_BYTES", "1")
        B1, res1, it1 = api.radiosity_solve(Fm, E, method=method, rel_tol=1e-10)
        monkeypatch.delenv("ILLUM_SPMM_CHUNK_BYTES")
        assert it == it1 and np.array_equal(res, res1) and np.array_equal(B, B1)
        # every dealing of row groups to the XCDs
        for xc in ("0", "1", "3", "16"):
            monkeypatch.setenv("ILLUM_RAD_XCD_CHUNK", xc)
            B1, res1, it1 = api.radiosity_solve(Fm, E, method=method, rel_tol=1e-10)
            assert it == it1 and np.array_equal(res, res1) and np.array_equal(B, B1), xc
        monkeypatch.delenv("ILLUM_RAD_XCD_CHUNK")
    keys = [t[0] for t in api.default_context().tuning()]
    assert any(k.startswith("direct_lanes[") for k in keys)
    assert any(k.startswith("spmm_per[fp64") for k in keys)
    assert any(k.startswith("spmm_per[fp16") for k in keys)
    assert any(k.startswith("spmm_xcd[fp64") for k in keys)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotuned_conduction_stream_layout_does_not_change_results(api, monkeypatch, precision):
    # Several depth buckets: the tuner times forked against sequential bucket
    # launches on a restored snapshot; the state after the run must equal a
    # run without tuning, and later runs reuse the choice.
    from illum_amd import configs
    rng = np.random.default_rng(8)
    nf = 2000
    grids = [configs.layer_grid(n) for n in (30, 45, 60, 75, 90)]
    gid = rng.integers(0, len(grids), nf).astype(np.int32)
    tis = np.exp(rng.normal(0, 0.5, nf))
    Q = rng.uniform(0, 600, (24, nf))
    out = []
    for tune in ("0", "1"):
        monkeypatch.setenv("ILLUM_NO_AUTOTUNE", "1" if tune == "0" else "0")
        th = api.ThermalModel(nf, grids, 200.0, 1.19177e6, 230.0, emissivity=0.95,
                              precision=precision, grid_of_facet=gid, ti_scale=tis)
        th.run(60.0, Q[:12])
        th.run(60.0, Q[12:])
        out.append((th.T, th.surface()))
        th.close()
    (Ta, (sa, fa)), (Tb, (sb, fb)) = out
    assert np.array_equal(Ta, Tb, equal_nan=True)
    assert np.array_equal(sa, sb) and np.array_equal(fa, fb)
    assert any(k.startswith("cn_streams[" + precision) for k, *_ in api.default_context().tuning())


@pytest.mark.parametrize("entries", [0, 1500])
def test_pool_overflow_reruns_sources_into_a_second_pool(api, monkeypatch, entries):
    # A first pool too small for the kept lists: the sources that did not
    # fit run again into a second pool, and the assembly is identical to one
    # whose pool held everything (CSC, CSR and the raw visibility sets).
    V, F = api.synth_ellipsoid(4, ncraters=40, seed=31)
    ic = api.IllumContext(V=V, F=F)
    ref = ic.compute_F(0.2, keep_visibility=True)
    rcsc, rcsr, rvis = ref.csc(), ref.csr(), ref.visibility_lists()
    assert ref.nnz > 2 * entries
    monkeypatch.setenv("ILLUM_FF_POOL_ENTRIES", str(entries))
    Fm = ic.compute_F(0.2, keep_visibility=True)
    for a, b in zip(Fm.csc(), rcsc):
        assert np.array_equal(a, b)
    for a, b in zip(Fm.csr(), rcsr):
        assert np.array_equal(a, b)
    for a, b in zip(Fm.visibility_lists(), rvis):
        assert np.array_equal(a, b)


def test_visibility_launch_orders_give_identical_assemblies(api, monkeypatch):
    # Sources dealt to the XCDs in chunks (default 8), as one range per XCD,
    # in plain order, or with rays walked unsorted: the assembly is the same.
    V, F = api.synth_ellipsoid(4, ncraters=60, seed=41)
    ic = api.IllumContext(V=V, F=F)
    ref = ic.compute_F(0.2)
    rcsc, rcsr = ref.csc(), ref.csr()
    for var, val in (("ILLUM_FF_XCD_CHUNK", "0"), ("ILLUM_FF_XCD_CHUNK", "1"),
                     ("ILLUM_FF_XCD_CHUNK", "5"), ("ILLUM_FF_RAY_SORT", "0"),
                     ("ILLUM_FF_RAY_CHUNK", "256"), ("ILLUM_FF_RAY_CHUNK", "1024")):
        monkeypatch.setenv(var, val)
        Fm = ic.compute_F(0.2)
        for a, b in zip(Fm.csc(), rcsc):
            assert np.array_equal(a, b), (var, val)
        for a, b in zip(Fm.csr(), rcsr):
            assert np.array_equal(a, b), (var, val)
        monkeypatch.delenv(var)


@pytest.mark.parametrize("npool", ["32", "1"])
def test_paged_visibility_set_gives_identical_assembly(api, monkeypatch, npool):
    # The visibility set kept in pool pages of 1,024 receivers instead of a
    # bitset over the body; with 1 pool page the sources whose visible
    # receivers span 2 pages (908 of 20,480 here) run out of pages and run
    # again with a bitset over their body.  CSC, CSR and the raw
    # visibility sets must equal the bitset assembly's.
    V, F = api.synth_ellipsoid(5, ncraters=80, seed=43)
    ic = api.IllumContext(V=V, F=F)
    monkeypatch.setenv("ILLUM_FF_PAGED", "0")
    ref = ic.compute_F(0.2, keep_visibility=True)
    rcsc, rcsr, rvis = ref.csc(), ref.csr(), ref.visibility_lists()
    rays = ref.rays()
    monkeypatch.setenv("ILLUM_FF_PAGED", "1")
    monkeypatch.setenv("ILLUM_FF_NPOOL", npool)
    ic.ctx.set_profiling(True)
    ic.ctx.reset_profile()
    Fm = ic.compute_F(0.2, keep_visibility=True)
    prof = ic.ctx.profile()
    ic.ctx.set_profiling(False)
    if npool == "1":
        assert prof.get("ff_visibility_rerun", (0, 0))[1] >= 1  # sources ran again
    for a, b in zip(Fm.csc(), rcsc):
        assert np.array_equal(a, b)
    for a, b in zip(Fm.csr(), rcsr):
        assert np.array_equal(a, b)
    for a, b in zip(Fm.visibility_lists(), rvis):
        assert np.array_equal(a, b)
    assert Fm.rays() == rays


def test_paged_work_count_with_sources_out_of_pool_pages(api, monkeypatch):
    # The instrumented pass (no lists written) with a 1-page pool: sources
    # out of pages store no counts and are counted, not run again.  Rays and
    # pair tests do not depend on how the visibility set is kept.
    V, F = api.synth_ellipsoid(5, ncraters=80, seed=43)
    ic = api.IllumContext(V=V, F=F)
    monkeypatch.setenv("ILLUM_FF_PAGED", "0")
    ref = ic.form_factor_work()
    monkeypatch.setenv("ILLUM_FF_PAGED", "1")
    monkeypatch.setenv("ILLUM_FF_NPOOL", "1")
    got = ic.form_factor_work()
    assert got["rays"] == ref["rays"] > 0
    assert got["pair_tests"] == ref["pair_tests"] > 0
    assert got["box_tests"] > 0 and got["tri_tests"] > 0


def test_body_whose_bitset_exceeds_the_lds_1310720_facets(api):
    # One body of 1,310,720 facets (icosphere subdivision 8): a bitset over
    # it (160 KB) leaves no LDS for the ray queue, so only the paged set can
    # run it.  Visibility sets of 4 x 16 sources bit-exact against the
    # oracle, their form-factor columns within 1e-12.
    from oracle import oracle as O
    V, F = api.synth_ellipsoid(8, ncraters=150, seed=3)
    nf = len(F)
    assert nf == 1310720
    ic = api.IllumContext(V=V, F=F, leaf_size=4)
    Fm = ic.compute_F(0.1, keep_visibility=True)
    vptr, vidx = Fm.visibility_lists()
    gcp, gri, gv = Fm.csc()
    g = O.geometry(V, F)
    worst = 0.0
    for j0 in (0, nf // 3, nf // 2 + 77, nf - 16):
        j1 = j0 + 16
        bits, _ = O.visibility(g, offset=1e-5, j0=j0, j1=j1)
        rows = np.unpackbits(bits.view(np.uint8), bitorder="little", axis=1)[:, :nf]
        for q, j in enumerate(range(j0, j1)):
            assert np.array_equal(vidx[vptr[j]:vptr[j + 1]], np.flatnonzero(rows[q])), j
        ocp, ori, ov = O.form_factors(g, bits, j0=j0, j1=j1, albedo=0.1)
        a, b = gcp[j0], gcp[j1]
        assert np.array_equal(gcp[j0:j1 + 1] - a, ocp)
        assert np.array_equal(gri[a:b], ori)
        if len(ov):
            worst = max(worst, float(np.max(np.abs(gv[a:b] - ov) / np.abs(ov))))
    assert worst <= 1e-12
    print(f"1,310,720 facets: nnz {Fm.nnz}, rays {Fm.rays()}, 64 sources bit-exact")
