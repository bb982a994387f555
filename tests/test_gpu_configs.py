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
