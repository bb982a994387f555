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
             vs mixed (FP32 sweeps + FP64 refinement)
  thermal    327,680 facets, 4,096 CN steps over 4 rotations, layer counts
             U{32..96}, thermal inertia lognormal (median 200, sigma 0.5),
             FP64 vs mixed
  coupled    radiosity + thermal time step on the 81,920-facet mesh
             (direct -> radiosity solve -> CN step), FP64 vs mixed

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
SUN_KM = 227390024.0          # illum_cli.cpp:337 (m -> km)
ROT_PERIOD_S = 5.342 * 3600   # Vesta, the reference's body (runjob.sh:3-4)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(V, F, offset, albedo_vec, nsample, nthreads):
    """CPU oracle (restatement of compute_F, own BVH) on a bounded sample of
    source facets, all host threads; returns pairs/s."""
    from oracle import oracle as O
    g = O.geometry(V, F)
    nf = len(F)
    blocks = 8
    per = max(1, nsample // blocks)
    t_total, pairs, rays = 0.0, 0, 0
    for b in range(blocks):
        j0 = (b * nf) // blocks
        t0 = time.perf_counter()
        bits, nr = O.visibility(g, offset=offset, j0=j0, j1=j0 + per, nthreads=nthreads)
        O.form_factors(g, bits, j0=j0, j1=j0 + per, albedo_vec=albedo_vec)
        t_total += time.perf_counter() - t0
        pairs += per * (nf - 1)
        rays += nr
    return pairs / t_total, dict(sources=per * blocks, seconds=t_total, rays=int(rays))


def hemisphere_suns(S, seed):
    rng = np.random.default_rng(seed)
    z = rng.uniform(0, 1, S)
    ph = rng.uniform(0, 2 * np.pi, S)
    r = np.sqrt(1 - z * z)
    return SUN_KM * np.stack([r * np.cos(ph), r * np.sin(ph), z])


def bench_horizons(api, illum, ctx, nphi=361, theta_eps=1e-3, offset=1e-5):
    import torch
    nf = illum.num_faces
    hz = torch.empty((nf, nphi), dtype=torch.float64, device="cuda")
    illum.make_horizons(nphi, theta_eps, offset, out=hz)  # warm-up
    ctx.synchronize()
    t = time.perf_counter()
    illum.make_horizons(nphi, theta_eps, offset, out=hz)
    ctx.synchronize()
    dt = time.perf_counter() - t
    rays = nf * nphi * 11
    return hz, {"seconds": dt, "rays_per_s": rays / dt, "rays": rays, "nphi": nphi}


def bench_radiosity(api, illum, ctx, Fm, hz_dev, S=512, tol=1e-8):
    import torch
    from illum_amd import _lib
    nf = illum.num_faces
    nphi = hz_dev.shape[1]
    sun = np.ascontiguousarray(hemisphere_suns(S, 1).T)  # S x 3 == 3 x S col-major
    disk = np.asfortranarray(api.fib_spiral(100))
    sun_d = torch.from_numpy(sun).cuda()
    disk_d = torch.from_numpy(np.ascontiguousarray(disk.T)).cuda()
    E = torch.empty((S, nf), dtype=torch.float64, device="cuda")  # nf x S col-major
    _lib.call("illum_direct_radiosity", illum.mesh.h, api._ptr(hz_dev), nphi, api._ptr(sun_d), S,
              api._ptr(disk_d), 100, api._ptr(E), _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_OUT_ON_DEVICE)
    ctx.synchronize()
    out = {"S": S, "rel_tol": tol, "E_mean": float(E.mean())}
    Bs = {}
    for method in ("fp64", "mixed"):
        B = torch.empty_like(E)
        resid = np.zeros(201)
        iters = api.C.c_int()
        m = {"fp64": _lib.ILLUM_RAD_JACOBI_FP64, "mixed": _lib.ILLUM_RAD_MIXED}[method]
        best = None
        for rep in range(2):
            ctx.synchronize()
            t = time.perf_counter()
            _lib.call("illum_radiosity_solve", Fm.h, api._ptr(E), S, m, tol, 200, api._ptr(B),
                      api._ptr(resid), api.C.byref(iters),
                      _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_OUT_ON_DEVICE)
            ctx.synchronize()
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        nres = int(np.count_nonzero(resid))
        out[method] = {"seconds": best, "products": iters.value,
                       "final_rel_resid": float(resid[max(0, nres - 1)])}
        Bs[method] = B
        log(f"radiosity {method}: {best:.3f}s, {iters.value} products, "
            f"resid {resid[max(0, nres - 1)]:.2e}")
    out["mixed_speedup"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    d = (Bs["mixed"] - Bs["fp64"]).abs().max().item()
    out["max_abs_diff_mixed_vs_fp64_W_m2"] = d
    out["B_max"] = Bs["fp64"].abs().max().item()
    return out


def bench_thermal(api, nsteps=4096, subdiv=7, seed=0):
    import torch
    V, F = api.synth_ellipsoid(subdiv, ncraters=150, seed=seed)
    nf = len(F)
    m = api.Mesh(api.default_context(), V=V, F=F)
    nrm = torch.from_numpy(m.geometry()["normal"].astype(np.float64)).cuda()
    m.close()
    rng = np.random.default_rng(seed)
    nzs = np.arange(32, 97)
    grids = [0.5 * (np.exp(np.linspace(0, 3, n)) - 1 + 0.01) / (np.exp(3) - 1 + 0.01) for n in nzs]
    gid = rng.integers(0, len(nzs), nf).astype(np.int32)
    tis = np.exp(rng.normal(0, 0.5, nf))
    rotations = 4
    dt = rotations * ROT_PERIOD_S / nsteps
    ang = 2 * np.pi * rotations * np.arange(1, nsteps + 1) / nsteps
    inc = np.deg2rad(20.0)
    sdir = np.stack([np.cos(ang) * np.cos(inc), np.sin(ang) * np.cos(inc), np.full(nsteps, np.sin(inc))])
    sd = torch.from_numpy(sdir).cuda()
    AU = 149.5978707e6
    S0 = 1372 * (AU / SUN_KM) ** 2 * (1 - 0.1)
    Q = (S0 * torch.clamp(sd.T @ nrm.T, min=0)).contiguous()  # nsteps x nf, device
    out = {"facets": nf, "steps": nsteps, "dt_s": dt, "layers": "U{32..96}",
           "cell_steps": int(np.sum(nzs[gid])) * nsteps}
    res = {}
    for prec in ("fp64", "mixed"):
        th = api.ThermalModel(nf, grids, 200.0, 1.19177e6, 200.0, emissivity=0.95,
                              precision=prec, grid_of_facet=gid, ti_scale=tis)
        th.run(dt, Q[:8])  # warm-up (also lays down Qprev)
        th2 = api.ThermalModel(nf, grids, 200.0, 1.19177e6, 200.0, emissivity=0.95,
                               precision=prec, grid_of_facet=gid, ti_scale=tis)
        th.close()
        api.default_context().synchronize()
        t = time.perf_counter()
        th2.run(dt, Q)
        api.default_context().synchronize()
        sec = time.perf_counter() - t
        Ts, _ = th2.surface()
        res[prec] = (th2.T, Ts)
        out[prec] = {"seconds": sec, "steps_per_s": nsteps / sec,
                     "cell_steps_per_s": out["cell_steps"] / sec}
        th2.close()
        log(f"thermal {prec}: {sec:.3f}s ({nsteps / sec:.0f} steps/s)")
    out["mixed_speedup"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    out["max_abs_Tsurf_diff_mixed_vs_fp64_K"] = float(np.max(np.abs(res["mixed"][1] - res["fp64"][1])))
    Tm, Tf = res["mixed"][0], res["fp64"][0]
    ok = ~np.isnan(Tf)
    out["max_abs_T_diff_mixed_vs_fp64_K"] = float(np.max(np.abs(Tm[ok] - Tf[ok])))
    # FP64 GPU vs the CPU oracle on a sample of facets
    from oracle import oracle as O
    samp = np.arange(0, nf, max(1, nf // 256))
    goff = np.cumsum([0] + [len(g) for g in grids])
    zc = np.concatenate(grids)
    Qh = Q[:, torch.from_numpy(samp).cuda()].cpu().numpy()
    T0 = np.full((97, len(samp)), 200.0)
    To, _, _, _ = O.thermal_run(goff, zc, np.full(len(zc), 200.0), np.full(len(zc), 1.19177e6),
                                T0, np.zeros(len(samp)), Qh, dt, 0.95, grid=gid[samp],
                                ti_scale=tis[samp])
    Tg = res["fp64"][0][:, samp]
    okk = ~np.isnan(Tg)
    out["max_abs_T_diff_fp64_vs_oracle_K"] = float(np.max(np.abs(Tg[okk] - To[okk])))
    Tgm = res["mixed"][0][:, samp]
    out["max_abs_Tsurf_diff_mixed_vs_oracle_K"] = float(np.max(np.abs(Tgm[0] - To[0])))
    return out


def bench_coupled(api, illum, ctx, Fm, hz_dev, nsteps=32):
    """radiosity + thermal time loop (illum_cli.cpp:403-456) on device buffers."""
    import torch
    from illum_amd import _lib
    nf = illum.num_faces
    nphi = hz_dev.shape[1]
    z = 0.5 * (np.exp(np.linspace(0, 3, 40)) - 1 + 0.01) / (np.exp(3) - 1 + 0.01)
    dt = ROT_PERIOD_S / 256
    disk_d = torch.from_numpy(np.ascontiguousarray(api.fib_spiral(100).T)).cuda()
    ang = 2 * np.pi * np.arange(nsteps) / 256
    suns = [torch.from_numpy(SUN_KM * np.array([np.cos(a) * 0.94, np.sin(a) * 0.94, 0.34])).cuda()
            for a in ang]
    out = {"steps": nsteps, "facets": nf, "layers": 40}
    finals = {}
    for prec in ("fp64", "mixed"):
        th = api.ThermalModel(nf, z, 200.0, 1.19177e6, 200.0, emissivity=0.95, precision=prec)
        E = torch.empty(nf, dtype=torch.float64, device="cuda")
        B = torch.empty_like(E)
        Fs = torch.empty_like(E)
        resid = np.zeros(101)
        it = api.C.c_int()
        method = _lib.ILLUM_RAD_JACOBI_FP64 if prec == "fp64" else _lib.ILLUM_RAD_MIXED
        flags = _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_OUT_ON_DEVICE
        ctx.synchronize()
        t = time.perf_counter()
        for k in range(nsteps):
            _lib.call("illum_direct_radiosity", illum.mesh.h, api._ptr(hz_dev), nphi, api._ptr(suns[k]),
                      1, api._ptr(disk_d), 100, api._ptr(E), flags)
            _lib.call("illum_thermal_get", th.h, None, None, api._ptr(Fs), _lib.ILLUM_OUT_ON_DEVICE)
            E.sub_(Fs)  # rad += -F_surf (illum_cli.cpp:416-418)
            _lib.call("illum_radiosity_solve", Fm.h, api._ptr(E), 1, method, 1e-8, 100, api._ptr(B),
                      api._ptr(resid), api.C.byref(it), flags)
            _lib.call("illum_thermal_step", th.h, dt, api._ptr(B), _lib.ILLUM_IN_ON_DEVICE)
        ctx.synchronize()
        api.default_context().synchronize()
        sec = time.perf_counter() - t
        finals[prec] = th.surface()[0]
        out[prec] = {"seconds": sec, "steps_per_s": nsteps / sec}
        th.close()
        log(f"coupled {prec}: {nsteps / sec:.1f} steps/s")
    out["mixed_speedup"] = out["fp64"]["seconds"] / out["mixed"]["seconds"]
    out["max_abs_Tsurf_diff_mixed_vs_fp64_K"] = float(np.max(np.abs(finals["mixed"] - finals["fp64"])))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--subdiv", type=int, default=6)
    ap.add_argument("--craters", type=int, default=150)
    ap.add_argument("--cpu-sample", type=int, default=512, help="source facets in the CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--thermal-steps", type=int, default=4096)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the form-factor path "
              "shards by source facet with no communication, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    from illum_amd import api
    offset = 1e-5
    t0 = time.perf_counter()
    V, F = api.synth_ellipsoid(args.subdiv, ncraters=args.craters, seed=0)
    nf = len(F)
    albedo = np.random.default_rng(0).uniform(0.05, 0.15, nf)
    ctx = api.default_context()
    illum = api.IllumContext(V=V, F=F, leaf_size=4, ctx=ctx)
    log(f"mesh {nf} facets, bvh {illum.bvh.info()}, setup {time.perf_counter() - t0:.1f}s")

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
    # Algorithmic bytes of ff_visibility per launch: every candidate pair
    # streams one 16-byte (centroid, radius) record; every source writes its
    # visibility row and its nonzero row (2 x N/8 bytes).
    words = ((nf + 511) // 512) * 16
    vis_bytes = pairs * 16 + nf * words * 4 * 2
    achieved = vis_bytes / (vis_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": "ff_visibility", "achieved": achieved,
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                "traffic": None, "avg_ms": vis_ms}

    secondary = {}
    if not args.quick:
        hz, secondary["horizons"] = bench_horizons(api, illum, ctx)
        log("horizons:", secondary["horizons"])
        secondary["radiosity"] = bench_radiosity(api, illum, ctx, Fm, hz)
        log("radiosity:", secondary["radiosity"])
        secondary["coupled"] = bench_coupled(api, illum, ctx, Fm, hz)
        log("coupled:", secondary["coupled"])
        del hz
        secondary["thermal"] = bench_thermal(api, nsteps=args.thermal_steps)
        log("thermal:", secondary["thermal"])

    cpu = None
    if not args.no_cpu:
        nth = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {args.cpu_sample} sources on {nth} threads ...")
        cv, info = cpu_baseline(V, F, offset, albedo, args.cpu_sample, nth)
        cpu = {"value": cv, "unit": "pairs/s", "cores": nth, "kind": "port",
               "sample": f"compute_F restated (oracle, own BVH) for {info['sources']} source facets "
                         f"in 8 evenly spaced contiguous blocks of the {nf}-facet mesh, "
                         f"{info['seconds']:.1f}s, {info['rays']} rays"}
        log(f"cpu {cv:.3e} pairs/s; gpu/cpu = {value / cv:.1f}x")

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
                   "facets": nf, "nnz": nnz, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "profile_ms": {k: v[0] / max(1, v[1]) for k, v in prof.items()},
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
