"""The device-resident coupled time loop (illum_timeloop_*, illum_cli.cpp:403-456)
against the same steps made call by call through the C ABI (direct
insolation, E -= F_surf, radiosity solve, CN step): identical temperatures,
radiosity and right-hand sides after every run, for graph replays of 1 and 8
steps, with other work and synchronisations between the runs, and for the
eager (uncaptured) form; both precisions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from illum_amd import api as A
    A.default_context()
    return A


@pytest.mark.parametrize("method,eager,direct", [
    ("fp64", False, "tuned"), ("mixed", False, "tuned"), ("fp64", True, "tuned"),
    ("mixed", True, "tuned"),
    # the two-pass direct insolation in replayed graphs: its queue must be
    # left empty by every launch for the next replay
    ("mixed", False, "two_pass")])
def test_timeloop_equals_call_by_call_steps(api, monkeypatch, method, eager, direct):
    from illum_amd import _lib, configs
    if direct == "two_pass":
        monkeypatch.setenv("ILLUM_DIRECT_LANES", "64")
    V, F = api.synth_ellipsoid(4, ncraters=40, seed=41)
    ic = api.IllumContext(V=V, F=F)
    Fm = ic.compute_F(0.1)
    nphi, nf = 65, len(F)
    c = api.default_context()
    hz = api.DeviceArray(c, (nf, nphi))
    ic.make_horizons(nphi, out=hz)
    disk = api.DeviceArray.from_host(c, np.ascontiguousarray(api.fib_spiral(100).T))
    z = configs.layer_grid(40)
    spr = 64
    dt = configs.ROT_PERIOD_S / spr
    runs = [3, 9, 8, 1, 11]  # remainders and whole 8-step graphs
    nsteps = sum(runs)
    sun = np.stack([configs.orbit_sun(k, spr) for k in range(nsteps)], 1)
    mk = lambda: api.ThermalModel(nf, z, configs.TI, configs.RHOC, 200.0, emissivity=configs.EMISS,
                                  precision=method)
    tha, thb = mk(), mk()
    loop = api.TimeLoop(Fm, ic, thb, hz, disk, method=method, rel_tol=1e-8, eager=eager)
    m = {"fp64": _lib.ILLUM_RAD_JACOBI_FP64, "mixed": _lib.ILLUM_RAD_MIXED}[method]
    dev = _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_OUT_ON_DEVICE | _lib.ILLUM_ASYNC
    E = api.DeviceArray(c, (nf,))
    B = api.DeviceArray(c, (nf,))
    sund = api.DeviceArray.from_host(c, np.ascontiguousarray(sun.T))
    resid = np.zeros(101)
    it = api.C.c_int()
    k = 0
    for r in runs:
        # call by call on model A
        for _ in range(r):
            p = api.C.c_void_p(sund.ptr.value + 24 * k)
            _lib.call("illum_direct_radiosity", ic.mesh.h, api._ptr(hz), nphi, p, 1,
                      api._ptr(disk), 100, api._ptr(E), dev)
            tha.get_radiosity(out=E, accumulate=True)
            _lib.call("illum_radiosity_solve", Fm.h, api._ptr(E), 1, m, 1e-8, 100, api._ptr(B),
                      api._ptr(resid), api.C.byref(it), dev)
            _lib.call("illum_thermal_step", tha.h, dt, api._ptr(B),
                      _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_ASYNC)
            k += 1
        # the loop on model B: the same suns, from step k - r
        loop.run(sun[:, k - r:k], dt)
        st = loop.stats()
        assert st["unconverged"] == 0 and st["max_resid"] <= 1e-8, st
        assert 1 <= st["max_products"] <= st["products"]
        Ta, Tb = tha.T, thb.T
        assert np.array_equal(Ta, Tb), f"temperatures differ after step {k}"
        El, Bl = loop.get()
        assert np.array_equal(El, E.to_host()) and np.array_equal(Bl, B.to_host())
        # unrelated work between runs (a solve on the same context)
        api.radiosity_solve(Fm, np.ones(nf), method=method)
    assert loop.stats()["steps"] == nsteps
    loop.close()
