"""Benchmark of the illum hot path on one MI355X.

Headline (one "step"): form-factor assembly (compute_F, illum.cpp:31-237) of a
seeded cratered triaxial ellipsoid -- semi-axes 13, 11, 9 km, icosahedron
subdivided 6 times (81,920 facets), 150 Gaussian craters (depth 0.05-0.3 km,
width 0.2-1.5 km), offset 1e-5, per-facet albedo U(0.05, 0.15) -- with the mesh
and BVH resident in HBM.  Metric: form-factor pairs/s = N(N-1) / t_step.

Secondary measurements on the same JSON line (skip with --quick):
  horizons   360 azimuths (nphi 361), theta_eps 1e-3 on the same mesh: rays/s
  radiosity  512 sun directions on the +z hemisphere, E from the direct
             insolation, (I - F) B = E to 1e-8 relative residual, FP64 Jacobi
             vs mixed (FP16-gathered corrections, FP64 residual refinement)
  thermal    327,680 facets, 4,096 CN steps over 4 rotations, layer counts
             U{32..96}, thermal inertia lognormal (median 200, sigma 0.5),
             FP64 vs mixed
  coupled    radiosity + thermal time step on the 81,920-facet mesh
             (direct -> radiosity solve -> CN step), FP64 vs mixed
  c5         64 bodies of 1,280-327,680 facets (5.43 M) through form factors,
             horizons, direct insolation under 32 suns per body, the mixed
             solve and 512 mixed CN steps in one launch sequence; per-stage
             seconds, and the smallest body checked against the oracle
  transfer   k-NN mesh-field transfer from the 81,920-facet mesh to the
             327,680-facet one (k = 5, FP64 brute force): pairs/s, checked
             against scipy's cKDTree on a target sample

Prints progress on stderr and ONE JSON line on stdout.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFS = 157.3
PEAK_FP64_TFS = 78.6
# Crank-Nicolson work per cell-step and the vector peak it is priced against:
# FP64 Thomas step (right-hand side 9, elimination 6 + 2 divisions, back
# substitution 2) on FP64 vector 78.6 TFLOP/s; mixed step (FP64 difference and
# update 3, FP32 coefficients, residual, factorisation, reciprocal and sweeps
# 16) on FP32 vector 157.3 TFLOP/s.
CN_FLOPS = {"fp64": (19, 78.6), "mixed": (19, 157.3)}
from illum_amd import configs  # noqa: E402  (seeded workloads of the configurations)

SUN_KM = configs.SUN_KM
ROT_PERIOD_S = configs.ROT_PERIOD_S


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(V, F, offset, albedo_vec, nsample, nthreads):
    """CPU oracle (restatement of compute_F, own BVH) on a bounded sample of
    source facets, all host threads; returns pairs/s and, per block, the
    oracle's visibility rows and CSC columns (the parity check's reference)."""
    from oracle import oracle as O
    g = O.geometry(V, F)
    nf = len(F)
    blocks = 8
    per = max(1, nsample // blocks)
    t_total, pairs, rays = 0.0, 0, 0
    ref = []
    for b in range(blocks):
        j0 = (b * nf) // blocks
        t0 = time.perf_counter()
        bits, nr = O.visibility(g, offset=offset, j0=j0, j1=j0 + per, nthreads=nthreads)
        csc = O.form_factors(g, bits, j0=j0, j1=j0 + per, albedo_vec=albedo_vec)
        t_total += time.perf_counter() - t0
        pairs += per * (nf - 1)
        rays += nr
        ref.append((j0, j0 + per, bits, csc))
    return pairs / t_total, dict(sources=per * blocks, seconds=t_total, rays=int(rays)), ref, g


def check_form_factors(Fm_vis, ref, nf):
    """GPU visibility sets and CSC columns against the oracle's blocks: rows
    that differ (bit-exact bar) and the largest relative form-factor error."""
    vptr, vidx = Fm_vis.visibility_lists()
    gcp, gri, gv = Fm_vis.csc()
    bad, worst, nrows = 0, 0.0, 0
    for j0, j1, bits, (ocp, ori, ov) in ref:
        rows = np.unpackbits(bits.view(np.uint8), bitorder="little", axis=1)[:, :nf]
        for q, j in enumerate(range(j0, j1)):
            bad += not np.array_equal(vidx[vptr[j]:vptr[j + 1]], np.flatnonzero(rows[q]))
        nrows += j1 - j0
        a, b = gcp[j0], gcp[j1]
        if not (np.array_equal(gcp[j0:j1 + 1] - a, ocp) and np.array_equal(gri[a:b], ori)):
            bad += 1
            continue
        if len(ov):
            worst = max(worst, float(np.max(np.abs(gv[a:b] - ov) / np.abs(ov))))
    return {"visibility_sets_checked": nrows, "visibility_sets_differing": int(bad),
            "form_factor_max_rel_err": worst}


def bench_horizons(api, illum, ctx, g, nphi=361, theta_eps=1e-3, offset=1e-5):
    from oracle import oracle as O
    nf = illum.num_faces
    hz = api.DeviceArray(ctx, (nf, nphi))  # nphi x nf column-major
    illum.make_horizons(nphi, theta_eps, offset, out=hz)  # warm-up
    ctx.reset_profile()
    ctx.set_profiling(True)
    t = time.perf_counter()
    illum.make_horizons(nphi, theta_eps, offset, out=hz)
    ctx.synchronize()
    dt = time.perf_counter() - t
    ctx.set_profiling(False)
    kms = ctx.profile().get("horizons", (float("nan"), 1))[0]
    rays = nf * nphi * 11
    # bit-exact check against the oracle on 8 blocks of 64 facets
    hh = hz.to_host()
    bad = 0
    for b in range(8):
        j0 = (b * nf) // 8
        ho, _ = O.horizons(g, nphi=nphi, theta_eps=theta_eps, offset=offset, j0=j0, j1=j0 + 64)
        bad += int(np.sum(np.any(hh[j0:j0 + 64] != ho.T, axis=1)))
    return hz, {"seconds": dt, "kernel_ms": kms, "rays_per_s": rays / (kms * 1e-3), "rays": rays,
                "nphi": nphi, "theta_eps": theta_eps, "facets_checked_vs_oracle": 512,
                "facets_differing": bad}


def bench_radiosity(api, illum, ctx, Fm, hz, S=512, tol=1e-8):
    from illum_amd import _lib
    nf = illum.num_faces
    nphi = hz.shape[1]
    sun = api.DeviceArray.from_host(ctx, np.ascontiguousarray(configs.hemisphere_suns(S, 1).T))
    disk = api.DeviceArray.from_host(ctx, np.ascontiguousarray(api.fib_spiral(100).T))
    E = api.DeviceArray(ctx, (S, nf))  # nf x S column-major
    dev = _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_OUT_ON_DEVICE | _lib.ILLUM_ASYNC
    _lib.call("illum_direct_radiosity", illum.mesh.h, api._ptr(hz), nphi, api._ptr(sun), S,
              api._ptr(disk), 100, api._ptr(E), dev)
    # direct insolation under the S suns, timed after the first call (which
    # autotunes the launch shape)
    ctx.synchronize()
    ctx.reset_profile()
    ctx.set_profiling(True)
    _lib.call("illum_direct_radiosity", illum.mesh.h, api._ptr(hz), nphi, api._ptr(sun), S,
              api._ptr(disk), 100, api._ptr(E), dev)
    ctx.synchronize()
    ctx.set_profiling(False)
    direct_ms = ctx.profile().get("direct", (float("nan"), 1))[0]
    out = {"S": S, "rel_tol": tol, "facets": nf, "nnz": Fm.nnz, "direct_ms": direct_ms}
    Bh = {}
    for method in ("fp64", "mixed"):
        B = api.DeviceArray(ctx, (S, nf))
        resid = np.zeros(201)
        iters = api.C.c_int()
        m = {"fp64": _lib.ILLUM_RAD_JACOBI_FP64, "mixed": _lib.ILLUM_RAD_MIXED}[method]
        best = None
        for rep in range(2):
            ctx.synchronize()
            ctx.reset_profile()
            ctx.set_profiling(True)
            t = time.perf_counter()
            _lib.call("illum_radiosity_solve", Fm.h, api._ptr(E), S, m, tol, 200, api._ptr(B),
                      api._ptr(resid), api.C.byref(iters), dev)
            ctx.synchronize()
            dt = time.perf_counter() - t
            ctx.set_profiling(False)
            best = dt if best is None else min(best, dt)
        nres = int(np.count_nonzero(resid))
        out[method] = {"seconds": best, "products": iters.value,
                       "final_rel_resid": float(resid[max(0, nres - 1)]),
                       "kernels_ms": {k: v[0] for k, v in ctx.profile().items()}}
        # per product against the rooflines: algorithmic FLOP 2 nnz S (FP64
        # FMAs in both methods) and the unique bytes streamed -- F once (12 B
        # per nonzero) plus per element the FP64 method's X, R and Y (24 B)
        # or the mixed method's FP16 correction in and out and FP32 remainder
        # in and out (10 B: S >= 128 at tolerances >= 1e-9, as here)
        sweep_ms = out[method]["kernels_ms"].get(f"radiosity_{method}_sweep")
        if sweep_ms and iters.value:
            pms = sweep_ms / iters.value
            flop = 2.0 * Fm.nnz * S
            ubytes = 12.0 * Fm.nnz + nf * S * (24 if method == "fp64" else 10)
            out[method]["per_product"] = {
                "ms": pms, "flop": flop, "tflops": flop / (pms * 1e-3) / 1e12,
                "frac_fp64_peak": flop / (pms * 1e-3) / 1e12 / PEAK_FP64_TFS,
                "unique_bytes": ubytes, "gbs": ubytes / (pms * 1e-3) / 1e9,
                "frac_hbm": ubytes / (pms * 1e-3) / 1e9 / PEAK_HBM_GBS}
        Bh[method] = B.to_host()
        B.close()
        log(f"radiosity {method}: {best:.3f}s, {iters.value} products, "
            f"resid {resid[max(0, nres - 1)]:.2e}")
    out["mixed_speedup"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    out["max_rel_diff_mixed_vs_fp64"] = float(np.max(np.abs(Bh["mixed"] - Bh["fp64"])) /
                                              np.max(np.abs(Bh["fp64"])))
    # FP64 relative residual of both answers, recomputed on the host from CSC
    import scipy.sparse as sps
    cp, ri, v = Fm.csc()
    A = sps.csc_matrix((v, ri, cp), shape=(nf, nf))
    Eh = E.to_host().T
    for method in ("fp64", "mixed"):
        Bm = Bh[method].T
        r = Eh - (Bm - A @ Bm)
        out[method]["host_check_rel_resid"] = float(np.max(np.linalg.norm(r, axis=0) /
                                                           np.linalg.norm(Eh, axis=0)))
    # B against the oracle's Gauss-Seidel (FP64, to 1e-13) for 2 of the suns
    from oracle import oracle as O
    for method in ("fp64", "mixed"):
        err = 0.0
        for s in (0, 311):
            Bo, _, _ = O.gauss_seidel(nf, cp, ri, v, Eh[:, s], max_sweeps=100, rel_tol=1e-13)
            err = max(err, float(np.linalg.norm(Bh[method][s] - Bo) / np.linalg.norm(Bo)))
        out[method]["rel_err_vs_oracle_gauss_seidel"] = err
    E.close()
    return out


def bench_thermal(api, nsteps=4096, subdiv=7, seed=0, chunk=256):
    # whole launches of `chunk` steps, at least one
    chunk = max(4, min(chunk, nsteps))
    nsteps = max(chunk, nsteps - nsteps % chunk)
    ctx = api.default_context()
    V, F = configs.mesh("C4", seed=seed)
    nf = len(F)
    m = api.Mesh(ctx, V=V, F=F)
    nrm = m.geometry()["normal"]
    m.close()
    grids, gid, tis = configs.c4_layers(nf, seed)
    nzs = np.array([len(g) for g in grids])
    rotations = 4
    spr = max(chunk, nsteps // rotations)
    dt = ROT_PERIOD_S / spr
    qh = [configs.insolation(nrm, c0, chunk, spr) for c0 in range(0, spr, chunk)]
    qd = [api.DeviceArray.from_host(ctx, q) for q in qh]
    out = {"facets": nf, "steps": nsteps, "dt_s": dt, "layers": "U{32..96}",
           "cell_steps": int(np.sum(nzs[gid])) * nsteps, "launches": nsteps // chunk}
    res = {}
    for prec in ("fp64", "mixed"):
        th = api.ThermalModel(nf, grids, 200.0, 1.19177e6, 200.0, emissivity=0.95,
                              precision=prec, grid_of_facet=gid, ti_scale=tis)
        th.run(dt, qd[0])  # warm-up (code load)
        th.close()
        th = api.ThermalModel(nf, grids, 200.0, 1.19177e6, 200.0, emissivity=0.95,
                              precision=prec, grid_of_facet=gid, ti_scale=tis)
        ctx.synchronize()
        ctx.reset_profile()
        ctx.set_profiling(True)
        t = time.perf_counter()
        for k in range(nsteps // chunk):
            th.run(dt, qd[k % len(qd)])
        ctx.synchronize()
        sec = time.perf_counter() - t
        ctx.set_profiling(False)
        kname = "cn_mixed" if prec == "mixed" else "cn_fp64"
        kms = ctx.profile().get(kname, (float("nan"), 1))[0]
        res[prec] = (th.T, th.surface()[0])
        th.close()
        # The column is register-resident (no per-step HBM stream), so the bound
        # is VALU: flops per cell-step counted from the kernels' arithmetic.
        fl, peak = CN_FLOPS[prec]
        tfs = out["cell_steps"] * fl / (kms * 1e-3) / 1e12
        out[prec] = {"seconds": sec, "kernel_ms": kms, "steps_per_s": nsteps / sec,
                     "cell_steps_per_s": out["cell_steps"] / (kms * 1e-3),
                     "flop_per_cell_step": fl, "TFLOPs": tfs, "valu_peak_TFLOPs": peak,
                     "valu_frac": tfs / peak}
        log(f"thermal {prec}: {sec:.3f}s wall, kernel {kms:.1f} ms ({nsteps / sec:.0f} steps/s)")
    out["mixed_speedup"] = out["fp64"]["kernel_ms"] / out["mixed"]["kernel_ms"]
    out["max_abs_Tsurf_diff_mixed_vs_fp64_K"] = float(np.max(np.abs(res["mixed"][1] - res["fp64"][1])))
    Tm, Tf = res["mixed"][0], res["fp64"][0]
    ok = ~np.isnan(Tf)
    out["max_abs_T_diff_mixed_vs_fp64_K"] = float(np.max(np.abs(Tm[ok] - Tf[ok])))
    # FP64 GPU and mixed GPU vs the CPU oracle on a sample of facets
    from oracle import oracle as O
    samp = np.arange(0, nf, max(1, nf // 256))
    goff = np.cumsum([0] + [len(g) for g in grids])
    zc = np.concatenate(grids)
    Qs = np.concatenate([qh[k % len(qh)][:, samp] for k in range(nsteps // chunk)])
    T0 = np.full((97, len(samp)), 200.0)
    To, _, _, _ = O.thermal_run(goff, zc, np.full(len(zc), 200.0), np.full(len(zc), 1.19177e6),
                                T0, np.zeros(len(samp)), Qs, dt, 0.95, grid=gid[samp],
                                ti_scale=tis[samp])
    Tg = Tf[:, samp]
    okk = ~np.isnan(Tg)
    out["max_abs_T_diff_fp64_vs_oracle_K"] = float(np.max(np.abs(Tg[okk] - To[okk])))
    out["max_abs_Tsurf_diff_mixed_vs_oracle_K"] = float(np.max(np.abs(Tm[0, samp] - To[0])))
    for d in qd:
        d.close()
    return out


def bench_c5(api, O):
    """C5 in one launch sequence (illum_amd/batch.py), wall-clock per stage
    with a device synchronisation around each.  The form-factor stage's rate
    counts the candidate pairs of every body, sum N_b (N_b - 1)."""
    from illum_amd import batch
    bodies = configs.c5_bodies()
    nb, S = len(bodies), 32
    suns = configs.c5_suns(nb, S)
    r = batch.run_bodies(bodies, suns, S, method="mixed", thermal="mixed")
    sizes = [len(F) for _, F in bodies]
    pairs = sum(n * (n - 1) for n in sizes)
    t = r["seconds"]
    out = {"bodies": nb, "facets": r["nf"], "facets_per_body": [min(sizes), max(sizes)],
           "suns_per_body": S, "cn_steps": r["steps"], "nnz": r["Fm"].nnz,
           "seconds": {k: round(v, 4) for k, v in t.items()},
           "form_factor_pairs_per_s": pairs / t["form_factors"],
           "radiosity_products": r["solve"][2], "radiosity_rel_resid": float(r["solve"][1][-1])}
    # the smallest body against the oracle run on it alone: visibility sets,
    # form factors and horizons
    b = sizes.index(min(sizes))
    Vb, Fb = bodies[b]
    lo, hi = int(r["offsets"][b]), int(r["offsets"][b + 1])
    g = O.geometry(Vb, Fb)
    bits, _ = O.visibility(g)
    ocp, ori, ov = O.form_factors(g, bits, albedo=0.1)
    cp, ri, v = r["Fm"].csc()
    a, z = cp[lo], cp[hi]
    same = bool(np.array_equal(cp[lo:hi + 1] - a, ocp) and np.array_equal(ri[a:z] - lo, ori))
    n = hi - lo
    hzb = r["hz"].to_host_flat(lo * r["nphi"], n * r["nphi"]).reshape(n, r["nphi"]).T
    ho, _ = O.horizons(g, nphi=r["nphi"])
    out["parity_body"] = {"body": b, "facets": n, "form_factor_pattern_equal": same,
                          "form_factor_max_rel_err": float(np.max(np.abs(v[a:z] - ov) / np.abs(ov)))
                          if same and len(ov) else None,
                          "horizons_equal": bool(np.array_equal(hzb, ho))}
    for k in ("hz", "E", "B"):
        r[k].close()
    r["th"].close()
    r["Fm"].close()
    return out


def bench_transfer(api):
    """k-NN transfer tables between the 81,920- and 327,680-facet meshes
    (build_mesh_transfer_tables.py's downsample query, k = 1, and its
    neighbour query, k = 5), hipEvent-free wall-clock around one call each."""
    from illum_amd import transfer
    from scipy.spatial import cKDTree
    Va, Fa = configs.mesh("C2")
    Vb, Fb = configs.mesh("C4")
    Pa, Pb = transfer.centroids(Va, Fa), transfer.centroids(Vb, Fb)
    transfer.knn(Pa, Pb[:4096], 5)  # loads the code object
    out = {"sources": len(Pa), "targets": len(Pb)}
    for k in (1, 5):
        t0 = time.perf_counter()
        idx, dist = transfer.knn(Pa, Pb, k)
        dt = time.perf_counter() - t0
        samp = np.arange(0, len(Pb), 97)
        dref, iref = cKDTree(Pa).query(Pb[samp], k)
        iref = iref.reshape(len(samp), k)
        dref = dref.reshape(len(samp), k)
        out[f"k{k}"] = {"seconds": dt, "pairs_per_s": len(Pa) * len(Pb) / dt,
                        "sample": len(samp),
                        "index_mismatches": int(np.sum(idx[samp] != iref)),
                        "max_abs_dist_err": float(np.max(np.abs(dist[samp] - dref)))}
    return out


def bench_coupled(api, illum, ctx, Fm, hz, nsteps=256):
    """The radiosity + thermal time loop (illum_cli.cpp:403-456) at one sun per
    step: direct -> E -= F_surf -> radiosity solve to 1e-8 -> CN step (40
    layers), device-resident, replayed from captured HIP graphs
    (illum_timeloop_*), FP64 (Jacobi + FP64 CN) vs mixed (FP16 corrections
    with FP64 residual + mixed CN).  The same steps made call by call through
    the C ABI are timed beside it."""
    from illum_amd import _lib
    nf = illum.num_faces
    nphi = hz.shape[1]
    spr = 256
    dt = ROT_PERIOD_S / spr
    z = configs.layer_grid(40)
    disk = api.DeviceArray.from_host(ctx, np.ascontiguousarray(api.fib_spiral(100).T))
    sun = np.stack([configs.orbit_sun(k, spr) for k in range(nsteps)], 1)
    sund = api.DeviceArray.from_host(ctx, np.ascontiguousarray(sun.T))
    out = {"steps": nsteps, "facets": nf, "layers": 40, "suns_per_step": 1, "rel_tol": 1e-8}
    finals = {}
    dev = _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_OUT_ON_DEVICE | _lib.ILLUM_ASYNC
    for prec in ("fp64", "mixed"):
        mk = lambda: api.ThermalModel(nf, z, configs.TI, configs.RHOC, 200.0,
                                      emissivity=configs.EMISS, precision=prec, ctx=ctx)
        th2 = mk()
        loop = api.TimeLoop(Fm, illum, th2, hz, disk, method=prec, rel_tol=1e-8)
        loop.run(sund, dt)  # first rotation: chooses the product budget, captures the graphs
        ctx.synchronize()
        t = time.perf_counter()
        loop.run(sund, dt)
        ctx.synchronize()
        sec = time.perf_counter() - t
        st = loop.stats()
        finals[prec] = th2.surface()[0]
        # call by call (direct, E -= F_surf, solve with host-checked stop, CN step)
        thc = mk()
        E = api.DeviceArray(ctx, (nf,))
        B = api.DeviceArray(ctx, (nf,))
        resid = np.zeros(101)
        it = api.C.c_int()
        m = {"fp64": _lib.ILLUM_RAD_JACOBI_FP64, "mixed": _lib.ILLUM_RAD_MIXED}[prec]
        ctx.synchronize()
        t = time.perf_counter()
        for k in range(nsteps):
            p = api.C.c_void_p(sund.ptr.value + 24 * k)
            _lib.call("illum_direct_radiosity", illum.mesh.h, api._ptr(hz), nphi, p, 1,
                      api._ptr(disk), 100, api._ptr(E), dev)
            thc.get_radiosity(out=E, accumulate=True)
            _lib.call("illum_radiosity_solve", Fm.h, api._ptr(E), 1, m, 1e-8, 100, api._ptr(B),
                      api._ptr(resid), api.C.byref(it), dev)
            _lib.call("illum_thermal_step", thc.h, dt, api._ptr(B),
                      _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_ASYNC)
        ctx.synchronize()
        sec_calls = time.perf_counter() - t
        out[prec] = {"seconds": sec, "steps_per_s": nsteps / sec, "products_budget": st["products"],
                     "max_products": st["max_products"], "unconverged": st["unconverged"],
                     "max_rel_resid": st["max_resid"],
                     "call_by_call_steps_per_s": nsteps / sec_calls}
        for x in (thc, E, B, loop, th2):
            x.close()
        log(f"coupled {prec}: {nsteps / sec:.1f} steps/s graph, {nsteps / sec_calls:.1f} call by call")
    out["mixed_speedup"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    out["max_abs_Tsurf_diff_mixed_vs_fp64_K"] = float(np.max(np.abs(finals["mixed"] - finals["fp64"])))
    disk.close()
    sund.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--subdiv", type=int, default=6)
    ap.add_argument("--craters", type=int, default=150)
    ap.add_argument("--cpu-sample", type=int, default=16384,
                    help="source facets in the CPU sample (about 10 s on 16 threads)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--leaf-size", type=int, default=0, help="BVH leaf size (0 = autotune)")
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--no-c5", action="store_true", help="skip the 64-body C5 leg")
    ap.add_argument("--thermal-steps", type=int, default=4096)
    ap.add_argument("--thermal-only", action="store_true",
                    help="only run the conduction leg and print its record")
    ap.add_argument("--c5-only", action="store_true", help="only the 64-body C5 leg")
    ap.add_argument("--horizons-only", action="store_true",
                    help="only the horizons leg (81,920 facets, nphi 361)")
    ap.add_argument("--transfer-only", action="store_true", help="only the k-NN transfer leg")
    ap.add_argument("--radiosity-only", action="store_true",
                    help="only the radiosity leg (81,920 facets, S = 512)")
    ap.add_argument("--coupled-only", action="store_true",
                    help="only run the coupled time-loop leg (81,920 facets) and print its record")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run one device-to-device copy of this many bytes (a known-size "
                         "kernel for scaling rocprof FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_bytes:
        from illum_amd import api as A
        c = A.default_context()
        # source written first (from the host) so every page holds data
        src = A.DeviceArray.from_host(c, np.ones(args.copy_bytes, np.uint8))
        dst = A.DeviceArray(c, (args.copy_bytes,), np.uint8)
        A._lib.call("illum_dev_copy", c.h, dst.ptr, src.ptr, args.copy_bytes, 2)
        c.synchronize()
        log(f"copied {args.copy_bytes} bytes device to device")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the form-factor path "
              "would shard by source facet (a contiguous block of CSC columns per GPU, no "
              "communication), which is not built here", file=sys.stderr)
        sys.exit(2)

    from illum_amd import api
    if args.thermal_only:
        print(json.dumps({"thermal": bench_thermal(api, nsteps=args.thermal_steps)}), flush=True)
        return
    if args.c5_only:
        from oracle import oracle as O
        print(json.dumps({"c5": bench_c5(api, O)}), flush=True)
        return
    if args.transfer_only:
        print(json.dumps({"transfer": bench_transfer(api)}), flush=True)
        return
    offset = 1e-5
    t0 = time.perf_counter()
    V, F = api.synth_ellipsoid(args.subdiv, ncraters=args.craters, seed=0)
    nf = len(F)
    albedo = configs.facet_albedo(nf)
    ctx = api.default_context()
    illum = api.IllumContext(V=V, F=F, leaf_size=args.leaf_size, ctx=ctx)
    log(f"mesh {nf} facets, bvh {illum.bvh.info()}, leaf autotune {illum.bvh.tuning()}, "
        f"setup {time.perf_counter() - t0:.1f}s")

    if args.horizons_only:
        from oracle import oracle as O
        hz, r = bench_horizons(api, illum, ctx, O.geometry(V, F))
        print(json.dumps({"horizons": r}), flush=True)
        return
    if args.coupled_only or args.radiosity_only:
        Fm = illum.compute_F(albedo, offset)
        hz = api.DeviceArray(ctx, (nf, 361))
        illum.make_horizons(361, 1e-3, offset, out=hz)
        if args.radiosity_only:
            print(json.dumps({"radiosity": bench_radiosity(api, illum, ctx, Fm, hz)}), flush=True)
        else:
            print(json.dumps({"coupled": bench_coupled(api, illum, ctx, Fm, hz)}), flush=True)
        return
    tune = illum.bvh.tuning()
    walk = ({2: "two-wide", 3: "wavefront packet (two-wide)", 4: "four-wide",
             5: "wavefront packet (four-wide)"}
            [min(tune, key=lambda x: x[2])[1]] if tune else "two-wide")
    for w in range(args.warmup):
        t = time.perf_counter()
        Fm = illum.compute_F(albedo, offset)
        log(f"warmup {w}: {time.perf_counter() - t:.3f}s nnz {Fm.nnz}")
        Fm.close()
    ctx.reset_profile()
    ctx.set_profiling(True)
    ctx.synchronize()
    times = []
    Fm = None
    for k in range(args.steps):
        if Fm is not None:
            Fm.close()
        ctx.synchronize()
        t = time.perf_counter()
        Fm = illum.compute_F(albedo, offset)
        ctx.synchronize()
        times.append(time.perf_counter() - t)
        log(f"step {k}: {times[-1]:.3f}s nnz {Fm.nnz}")
    ctx.set_profiling(False)
    prof = ctx.profile()
    log("profile (ms, launches):", prof)
    nnz = Fm.nnz
    t_step = float(np.mean(times))
    pairs = nf * (nf - 1)
    value = pairs / t_step

    vis_ms, vis_n = prof.get("ff_visibility", (float("nan"), 1))
    vis_ms /= max(1, vis_n)
    rays = Fm.rays()
    # VALU roofline of ff_visibility from algorithmic work, FP32 FLOP per
    # launch: 8 per candidate pair (j, i != j), all N(N-1) of them (the facing
    # cull of illum.cpp:56-68 as written, before any cluster pass), plus 12 per
    # ray/box slab test and 40 per ray/triangle test counted per ray along its
    # own near-first walk of the same tree (the per-lane two-wide walk's
    # counters, in a separate instrumented launch; they equal exactly the
    # oracle's walk of the exported tree, oracle.walk_work, as
    # test_per_ray_walk_work_equals_the_oracle_walk_of_the_same_tree checks)
    # -- not the wavefront packet walk's union of nodes, which is how the
    # kernel spends its time.
    os.environ["ILLUM_FF_TRAVERSAL"] = "wide2"
    st = illum.form_factor_work(offset)
    del os.environ["ILLUM_FF_TRAVERSAL"]
    st_packet = illum.form_factor_work(offset)  # the default walk, for the record
    pairs = nf * (nf - 1)
    flops = 8 * pairs + 12 * st["box_tests"] + 40 * st["tri_tests"]
    tflops = flops / (vis_ms * 1e-3) / 1e12
    # HBM, secondary: algorithmic bytes of the launch are the facet records read
    # once (centroid+radius and normal 32 B, area 8 B: 40 B per facet), the
    # per-source count and pool offset (12 B per facet) and, per kept entry,
    # its 4 B pool index and 4 B CSR row-count update.
    vis_bytes = nf * 52 + nnz * 8
    achieved_gbs = vis_bytes / (vis_ms * 1e-3) / 1e9
    traffic = None
    pmc_path = os.path.join(ROOT, "profiles", f"ff_visibility_{nf}_pmc.json")
    if os.path.exists(pmc_path):
        pmc = json.load(open(pmc_path))
        traffic = (pmc["fetch_kib"] * pmc["fetch_scale"] + pmc["write_kib"] * pmc["write_scale"]) * 1024
    roofline = {"bound": "valu", "kernel": "ff_visibility", "achieved": tflops,
                "peak": PEAK_FP32_TFS, "unit": "TFLOP/s", "frac": tflops / PEAK_FP32_TFS,
                "traffic": traffic, "avg_ms": vis_ms, "flop": flops,
                "work": {"pairs": pairs, "rays": rays, "box_tests_per_lane_walk": st["box_tests"],
                         "tri_tests_per_lane_walk": st["tri_tests"],
                         "packet_walk_counters": st_packet},
                "rays_per_s": rays / (vis_ms * 1e-3),
                "hbm": {"algorithmic_bytes": vis_bytes, "achieved": achieved_gbs,
                        "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved_gbs / PEAK_HBM_GBS,
                        "traffic": traffic}}

    secondary = {}
    if not args.quick:
        from oracle import oracle as O
        hz, secondary["horizons"] = bench_horizons(api, illum, ctx, O.geometry(V, F))
        log("horizons:", secondary["horizons"])
        secondary["radiosity"] = bench_radiosity(api, illum, ctx, Fm, hz)
        log("radiosity:", secondary["radiosity"])
        secondary["coupled"] = bench_coupled(api, illum, ctx, Fm, hz)
        log("coupled:", secondary["coupled"])
        hz.close()
        secondary["thermal"] = bench_thermal(api, nsteps=args.thermal_steps)
        log("thermal:", secondary["thermal"])
        if not args.no_c5:
            secondary["c5"] = bench_c5(api, O)
            log("c5:", secondary["c5"])
        secondary["transfer"] = bench_transfer(api)
        log("transfer:", secondary["transfer"])

    cpu = None
    parity = {}
    if not args.no_cpu:
        nth = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {args.cpu_sample} sources on {nth} threads ...")
        cv, info, ref, _ = cpu_baseline(V, F, offset, albedo, args.cpu_sample, nth)
        cpu = {"value": cv, "unit": "pairs/s", "cores": nth, "kind": "port",
               "cpu_model": cpu_model(), "nproc": os.cpu_count(),
               "sample": f"compute_F restated (oracle, own BVH) for {info['sources']} source facets "
                         f"in 8 evenly spaced contiguous blocks of the {nf}-facet mesh, "
                         f"{info['seconds']:.1f}s, {info['rays']} rays"}
        log(f"cpu {cv:.3e} pairs/s; gpu/cpu = {value / cv:.1f}x")
        # the same sources' visibility sets and columns from the GPU, assembled
        # again outside the timed steps with the raw sets kept
        Fv = illum.compute_F(albedo, offset, keep_visibility=True)
        parity["form_factors"] = check_form_factors(Fv, ref, nf)
        Fv.close()
        log("parity:", parity["form_factors"])
    for leg in ("horizons", "radiosity", "thermal"):
        if leg in secondary:
            r = secondary[leg]
            if leg == "horizons":
                parity[leg] = {k: r[k] for k in ("facets_checked_vs_oracle", "facets_differing")}
            elif leg == "radiosity":
                parity[leg] = {m: {k: r[m][k] for k in ("final_rel_resid", "host_check_rel_resid",
                                                         "rel_err_vs_oracle_gauss_seidel")}
                               for m in ("fp64", "mixed")}
            else:
                parity[leg] = {k: r[k] for k in ("max_abs_T_diff_fp64_vs_oracle_K",
                                                  "max_abs_Tsurf_diff_mixed_vs_oracle_K")}

    out = {
        "metric": "form-factor pairs/s",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 visibility, f64 form factors",
        "data": "synthetic",
        "config": {"workload": f"compute_F on seeded cratered ellipsoid, {nf} facets "
                               f"(icosphere subdiv {args.subdiv}, {args.craters} craters), "
                               "per-facet albedo U(0.05,0.15), CSC+CSR assembly",
                   "facets": nf, "nnz": nnz, "parallelism": "single GPU",
                   "bvh": dict(illum.bvh.info(), walk=walk)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "profile_ms": {k: v[0] / max(1, v[1]) for k, v in prof.items()},
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
