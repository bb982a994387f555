"""Many bodies of uneven size through the whole path in one launch sequence:
form factors, horizon maps, direct insolation under each body's own suns,
the radiosity solve and one rotation of Crank-Nicolson steps, each stage one
launch (or one launch per time step) over the concatenated bodies.

The reference handles one body per process (illum_cli.cpp:270-510); here the
bodies share every launch: visibility, form factors and horizons stay within a
facet's own body (F is block-diagonal), and the solve and conduction run over
all facets at once.
"""
import ctypes as C
import time

import numpy as np

from . import _lib, api


def run_bodies(bodies, suns, S, nphi=361, albedo=0.1, method="mixed", thermal="mixed",
               steps=512, rel_tol=1e-8, layers=40, leaf_size=3, ctx=None, keep_visibility=False):
    """bodies: [(V, F)]; suns: 3 x (S nbody) km, body b's in columns
    b S .. b S + S - 1.  The conduction run covers one rotation of `steps`
    steps; step k is driven by the scattered radiosity of sun k S // steps
    (its B column, read on the device).  Returns a dict with the context,
    matrix, device arrays hz (nphi x N), E and B (N x S), the thermal model,
    the body offsets and per-stage seconds.  leaf_size 3: the horizon walks,
    most of the batch's time, are faster on small leaves, and the batch is
    too large to build a tree per candidate leaf size for the autotuner."""
    ctx = ctx or api.default_context()
    t = {}

    def stage(name, fn):
        ctx.synchronize()
        t0 = time.perf_counter()
        r = fn()
        ctx.synchronize()
        t[name] = time.perf_counter() - t0
        return r

    V, F, off = api.concat_bodies(bodies)
    nf = len(F)
    nb = len(bodies)
    ic = stage("setup", lambda: api.IllumContext(V=V, F=F, bodies=off, leaf_size=leaf_size, ctx=ctx))
    Fm = stage("form_factors", lambda: ic.compute_F(albedo, keep_visibility=keep_visibility))
    hz = api.DeviceArray(ctx, (nf, nphi))
    stage("horizons", lambda: ic.make_horizons(nphi, out=hz))
    dev = _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_OUT_ON_DEVICE | _lib.ILLUM_ASYNC
    sun_d = api.DeviceArray.from_host(ctx, np.ascontiguousarray(np.asarray(suns, np.float64).T))
    assert sun_d.shape == (S * nb, 3)
    disk = api.DeviceArray.from_host(ctx, np.ascontiguousarray(api.fib_spiral(100).T))
    E = api.DeviceArray(ctx, (S, nf))
    stage("direct", lambda: _lib.call("illum_direct_radiosity", ic.mesh.h, api._ptr(hz), nphi,
                                      api._ptr(sun_d), S, api._ptr(disk), 100, api._ptr(E),
                                      dev | _lib.ILLUM_SUN_PER_BODY))
    B = api.DeviceArray(ctx, (S, nf))
    solve = stage("radiosity", lambda: api.radiosity_solve(Fm, E, method=method, rel_tol=rel_tol,
                                                           max_iter=200, B=B))
    from .configs import EMISS, RHOC, ROT_PERIOD_S, TI, layer_grid
    th = api.ThermalModel(nf, layer_grid(layers), TI, RHOC, 200.0, emissivity=EMISS,
                          precision=thermal, ctx=ctx)
    dt = ROT_PERIOD_S / steps
    base = B.ptr.value

    def conduction():
        # step k is driven by B column (k S) // steps: the steps of one column
        # run as one launch sequence with that column held
        k = 0
        while k < steps:
            s = (k * S) // steps
            k1 = k
            while k1 < steps and (k1 * S) // steps == s:
                k1 += 1
            _lib.call("illum_thermal_run", th.h, dt, C.c_void_p(base + 8 * nf * s), k1 - k,
                      _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_ASYNC | _lib.ILLUM_Q_HELD)
            k = k1

    stage("conduction", conduction)
    sun_d.close()
    disk.close()
    return dict(ic=ic, Fm=Fm, hz=hz, E=E, B=B, th=th, offsets=off, seconds=t, solve=solve,
                nf=nf, S=S, nphi=nphi, steps=steps, dt=dt)
