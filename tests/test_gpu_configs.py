"""GPU parity at the measured configurations' full sizes (illum_amd/configs.py),
against the oracle on samples it finishes in seconds.

  C2  81,920 facets: visibility sets bit-exact for 16,384 sources (8 blocks of
      2,048), their form-factor columns within 1e-12 relative, CSR equal to
      CSC transposed over all nonzeros, horizons at nphi 361 bit-exact on 512
      facets (illum.cpp:42-234, 326-363)
  C3  81,920 facets x 512 suns: direct insolation within 1e-12 of the oracle
      on a facet sample; both solve methods at <= 1e-8 relative residual,
      recomputed in FP64 on the host; B within 2e-8 (2-norm, relative) of the
      oracle's Gauss-Seidel for 4 suns (illum_cli.cpp:425-440)
  C4  327,680 facets x 4,096 CN steps, ragged 32-96 layers, lognormal thermal
      inertia: FP64 within 1e-9 K and the mixed surface temperature within
      1e-3 K of the oracle on 256 facets (thermal.hpp:65-84, conductionQ.f)
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


@pytest.fixture(scope="module")
def c2(api):
    from illum_amd import configs
    V, F = configs.mesh("C2")
    albedo = configs.facet_albedo(len(F))
    ic = api.IllumContext(V=V, F=F, leaf_size=4)
    Fm = ic.compute_F(albedo, keep_visibility=True)
    g = O.geometry(V, F)
    return dict(V=V, F=F, albedo=albedo, ic=ic, Fm=Fm, g=g)


def _blocks(nf, nblock=8, per=2048):
    return [((b * nf) // nblock, (b * nf) // nblock + per) for b in range(nblock)]


def test_c2_visibility_and_form_factors_81920(api, c2):
    nf = len(c2["F"])
    assert nf == 81920
    Fm, g, albedo = c2["Fm"], c2["g"], c2["albedo"]
    vptr, vidx = Fm.visibility_lists()
    gcp, gri, gv = Fm.csc()
    worst = 0.0
    for j0, j1 in _blocks(nf):
        bits, _ = O.visibility(g, offset=1e-5, j0=j0, j1=j1)
        rows = np.unpackbits(bits.view(np.uint8), bitorder="little", axis=1)[:, :nf]
        for q, j in enumerate(range(j0, j1)):
            want = np.flatnonzero(rows[q])
            got = vidx[vptr[j]:vptr[j + 1]]
            assert np.array_equal(got, want), f"visibility set of source {j}"
        ocp, ori, ov = O.form_factors(g, bits, j0=j0, j1=j1, albedo_vec=albedo)
        a, b = gcp[j0], gcp[j1]
        assert np.array_equal(gcp[j0:j1 + 1] - a, ocp), f"columns {j0}..{j1}"
        assert np.array_equal(gri[a:b], ori)
        worst = max(worst, float(np.max(np.abs(gv[a:b] - ov) / np.abs(ov))) if len(ov) else 0.0)
    print(f"C2: 16,384 visibility sets bit-exact, form factors max rel err {worst:.2e}")
    assert worst <= 1e-12
    # CSR = CSC transposed, every nonzero
    import scipy.sparse as sps
    A = sps.csc_matrix((gv, gri, gcp), shape=(nf, nf)).tocsr()
    A.sort_indices()
    rp, ci, rv = Fm.csr()
    assert np.array_equal(rp, A.indptr) and np.array_equal(ci, A.indices)
    assert np.array_equal(rv, A.data)


def test_c2_horizons_nphi_361(api, c2):
    nf = len(c2["F"])
    hz = np.asarray(c2["ic"].make_horizons(nphi=361))
    for j0, j1 in _blocks(nf, 8, 64):
        ho, nr = O.horizons(c2["g"], nphi=361, theta_eps=1e-3, offset=1e-5, j0=j0, j1=j1)
        assert nr == (j1 - j0) * 361 * 11
        assert np.array_equal(hz[:, j0:j1], ho), f"horizons of facets {j0}..{j1}"
    c2["hz"] = hz


def test_c3_512_suns_both_methods(api, c2):
    from illum_amd import _lib, configs
    nf = len(c2["F"])
    ic, Fm, g = c2["ic"], c2["Fm"], c2["g"]
    hz = c2.get("hz")
    if hz is None:
        hz = np.asarray(ic.make_horizons(nphi=361))
    S = 512
    sun = configs.hemisphere_suns(S, 1)
    disk = api.fib_spiral(100)
    E = ic.get_direct_radiosity(sun, disk)
    assert E.shape == (nf, S)
    # direct insolation against the oracle on a facet sample, 8 suns
    samp = np.arange(0, nf, 331)
    gs = {k: g[k][samp] for k in ("centroid", "frame", "tri")}
    Eo = O.direct(gs, hz[:, samp], sun[:, :8], disk)
    np.testing.assert_allclose(E[samp, :8], Eo, rtol=1e-12, atol=0)
    import scipy.sparse as sps
    cp, ri, v = Fm.csc()
    A = sps.csc_matrix((v, ri, cp), shape=(nf, nf))
    out = {}
    for method in ("fp64", "mixed"):
        B, res, it = api.radiosity_solve(Fm, E, method=method, rel_tol=1e-8, max_iter=200)
        r = E - (B - A @ B)
        host = float(np.max(np.linalg.norm(r, axis=0) / np.linalg.norm(E, axis=0)))
        print(f"C3 {method}: {it} products, reported {res[-1]:.2e}, host FP64 residual {host:.2e}")
        assert res[-1] <= 1e-8 and host <= 1e-8
        out[method] = B
    for s in (0, 101, 255, 511):
        Bo, _, _ = O.gauss_seidel(nf, cp, ri, v, E[:, s], max_sweeps=100, rel_tol=1e-13)
        for method, B in out.items():
            err = np.linalg.norm(B[:, s] - Bo) / np.linalg.norm(Bo)
            assert err <= 2e-8, (method, s, err)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("mixed", 1e-3)])
def test_c4_327680_facets_4096_steps(api, precision, tol):
    from illum_amd import configs
    V, F = configs.mesh("C4")
    nf = len(F)
    assert nf == 327680
    ctx = api.default_context()
    m = api.Mesh(ctx, V=V, F=F)
    nrm = m.geometry()["normal"]
    m.close()
    grids, gid, tis = configs.c4_layers(nf)
    nsteps, chunk, spr = 4096, 256, 1024
    dt = configs.ROT_PERIOD_S / spr
    qh = [configs.insolation(nrm, c0, chunk, spr) for c0 in range(0, spr, chunk)]
    qd = [api.DeviceArray.from_host(ctx, q) for q in qh]
    th = api.ThermalModel(nf, grids, configs.TI, configs.RHOC, 200.0, emissivity=configs.EMISS,
                          precision=precision, grid_of_facet=gid, ti_scale=tis)
    for k in range(nsteps // chunk):
        th.run(dt, qd[k % len(qd)])
    T = th.T
    th.close()
    for d in qd:
        d.close()
    samp = np.arange(0, nf, nf // 256)
    goff = np.cumsum([0] + [len(gr) for gr in grids])
    zc = np.concatenate(grids)
    Qs = np.concatenate([qh[k % len(qh)][:, samp] for k in range(nsteps // chunk)])
    T0 = np.full((97, len(samp)), 200.0)
    To, _, _, _ = O.thermal_run(goff, zc, np.full(len(zc), configs.TI), np.full(len(zc), configs.RHOC),
                                T0, np.zeros(len(samp)), Qs, dt, configs.EMISS, grid=gid[samp],
                                ti_scale=tis[samp])
    Tg = T[:, samp]
    ok = ~np.isnan(Tg)
    # rows past a facet's layer count are NaN in the GPU's (nzmax+1) x N copy
    nz = np.array([len(grids[k]) for k in gid[samp]])
    assert np.array_equal(ok, np.arange(97)[:, None] <= nz[None, :])
    dT = float(np.max(np.abs(Tg[ok] - To[ok])))
    dTs = float(np.max(np.abs(Tg[0] - To[0])))
    print(f"C4 {precision}: max |T - oracle| {dT:.2e} K, surface {dTs:.2e} K")
    if precision == "fp64":
        assert dT <= tol
    else:
        assert dTs <= tol


@pytest.mark.parametrize("method", ["mixed"])
def test_c5_64_bodies_one_launch_sequence(api, method):
    """C5: 64 bodies (5.4 M facets, 1,280 to 327,680 each) through form
    factors, horizons (nphi 361), direct insolation under each body's 32 suns,
    the radiosity solve (S = 32) and 512 CN steps, one launch sequence
    (illum_amd/batch.py).  Per body, against the oracle run on that body
    alone, for two 1,280-facet bodies, a 5,120-facet body and a 20,480-facet
    body (horizons and temperatures on a facet sample there): form-factor
    columns within 1e-12, horizons bit-exact, E within 1e-12, B within 2e-8
    of Gauss-Seidel, surface temperatures within 1e-3 K (FP64 conduction
    from the oracle's own B)."""
    from illum_amd import batch, configs
    bodies = configs.c5_bodies()
    nb, S = len(bodies), 32
    suns = configs.c5_suns(nb, S)
    r = batch.run_bodies(bodies, suns, S, method=method, thermal=method)
    nf, off = r["nf"], r["offsets"]
    assert nb == 64 and nf == 5434880
    print("C5 stage seconds:", {k: round(v, 3) for k, v in r["seconds"].items()},
          "products", r["solve"][2], "residual", r["solve"][1][-1])
    assert r["solve"][1][-1] <= 1e-8
    cp, ri, v = r["Fm"].csc()
    disk = api.fib_spiral(100)
    sizes = [len(F) for _, F in bodies]
    picks = [sizes.index(1280), len(sizes) - 1 - sizes[::-1].index(1280), sizes.index(5120),
             sizes.index(20480)]
    errs = {}
    for b in picks:
        Vb, Fb = bodies[b]
        lo, hi = int(off[b]), int(off[b + 1])
        n = hi - lo
        g = O.geometry(Vb, Fb)
        bits, _ = O.visibility(g)
        ocp, ori, ov = O.form_factors(g, bits, albedo=0.1)
        a, z = cp[lo], cp[hi]
        assert np.array_equal(cp[lo:hi + 1] - a, ocp), f"body {b} columns"
        assert np.array_equal(ri[a:z] - lo, ori), f"body {b} rows"
        ff_err = float(np.max(np.abs(v[a:z] - ov) / np.abs(ov))) if len(ov) else 0.0
        samp = np.arange(n) if n <= 5120 else np.arange(0, n, n // 256)
        hz = r["hz"].to_host_flat(lo * 361, n * 361).reshape(n, 361).T
        if n <= 5120:
            ho, _ = O.horizons(g, nphi=361)
        else:
            ho = np.concatenate([O.horizons(g, nphi=361, j0=j, j1=j + 1)[0] for j in samp], 1)
        assert np.array_equal(hz[:, samp], ho), f"body {b} horizons"
        sb = suns[:, b * S:(b + 1) * S]
        E = np.stack([r["E"].to_host_flat(s * nf + lo, n) for s in range(S)], 1)
        Eo = O.direct(g, hz, sb, disk)
        e_err = float(np.max(np.abs(E - Eo) / np.maximum(np.abs(Eo), 1e-300)))
        assert np.allclose(E, Eo, rtol=1e-12, atol=0), f"body {b} direct"
        B = np.stack([r["B"].to_host_flat(s * nf + lo, n) for s in range(S)], 1)
        Bo = np.stack([O.gauss_seidel(n, ocp, ori, ov, Eo[:, s], max_sweeps=200,
                                      rel_tol=1e-13)[0] for s in range(S)], 1)
        b_err = float(np.max(np.linalg.norm(B - Bo, axis=0) / np.linalg.norm(Bo, axis=0)))
        assert b_err <= 2e-8, f"body {b} radiosity {b_err}"
        # conduction of the sampled facets under the oracle's B
        steps = r["steps"]
        Q = np.stack([Bo[samp, (k * S) // steps] for k in range(steps)])
        z = configs.layer_grid(40)
        To, _, _, _ = O.thermal_run([0, 40], z, np.full(40, configs.TI), np.full(40, configs.RHOC),
                                    np.full((41, len(samp)), 200.0), np.zeros(len(samp)), Q,
                                    r["dt"], configs.EMISS)
        Ts = r["th"].surface()[0][lo + samp]
        t_err = float(np.max(np.abs(Ts - To[0])))
        assert t_err <= 1e-3, f"body {b} surface temperature {t_err}"
        errs[b] = dict(facets=n, ff=ff_err, E=e_err, B=b_err, Tsurf_K=t_err)
    print("C5 per-body error vs oracle:", errs)
