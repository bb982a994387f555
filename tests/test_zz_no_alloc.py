"""illum_radiosity_solve_noalloc (no device allocation: scratch from a fixed
per-GPU array) against illum_radiosity_solve on the same inputs, through the
C ABI: identical B, residual history and product count, on two calls in a
row (the second smaller, so it runs over what the first left in the array),
in both methods; the ring mode (S >= 128) and B mode; refusals before any
work for host buffers, max_iter above the bound and problems above the
array's size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from illum_amd import api as A
    A.default_context()
    return A


def _solve(api, entry, Fm, Ed, Bd, S, method, tol, max_iter):
    from illum_amd import _lib
    m = {"fp64": _lib.ILLUM_RAD_JACOBI_FP64, "mixed": _lib.ILLUM_RAD_MIXED}[method]
    resid = np.zeros(max_iter + 1)
    it = api.C.c_int()
    _lib.call(entry, Fm.h, api._ptr(Ed), S, m, tol, max_iter, api._ptr(Bd), api._ptr(resid),
              api.C.byref(it), _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_OUT_ON_DEVICE)
    return Bd.to_host(), resid[:it.value].copy(), it.value


@pytest.mark.parametrize("method", ["fp64", "mixed"])
def test_noalloc_solve_equals_workspace_solve(api, method):
    ctx = api.default_context()
    cases = []
    # first call: 5,120 facets, S = 130 (ring mode for the mixed method);
    # second: 1,280 facets, S = 3 (B mode)
    for subdiv, S, seed in ((4, 130, 1), (3, 3, 2)):
        V, F = api.synth_ellipsoid(subdiv, ncraters=40, seed=seed)
        ic = api.IllumContext(V=V, F=F)
        Fm = ic.compute_F(0.5)
        nf = len(F)
        E = np.random.default_rng(seed).uniform(0, 400, (S, nf))  # nf x S column-major
        cases.append((ic, Fm, S, api.DeviceArray.from_host(ctx, E)))
    for ic, Fm, S, Ed in cases:
        nf = Fm.n
        B1 = api.DeviceArray(ctx, (S, nf))
        B2 = api.DeviceArray(ctx, (S, nf))
        a = _solve(api, "illum_radiosity_solve", Fm, Ed, B1, S, method, 1e-11, 200)
        b = _solve(api, "illum_radiosity_solve_noalloc", Fm, Ed, B2, S, method, 1e-11, 200)
        assert a[2] == b[2] and a[2] >= 3
        assert np.array_equal(a[1], b[1]) and a[1][-1] <= 1e-11
        assert np.array_equal(a[0], b[0])


def test_noalloc_solve_refusals(api):
    from illum_amd import _lib
    ctx = api.default_context()
    V, F = api.synth_ellipsoid(3, ncraters=40, seed=3)
    Fm = api.IllumContext(V=V, F=F).compute_F(0.3)
    nf = Fm.n
    E = np.random.default_rng(0).uniform(0, 400, nf)
    Ed = api.DeviceArray.from_host(ctx, E)
    Bd = api.DeviceArray(ctx, (nf,))
    resid = np.zeros(400)
    it = api.C.c_int()
    dev = _lib.ILLUM_IN_ON_DEVICE | _lib.ILLUM_OUT_ON_DEVICE
    m = _lib.ILLUM_RAD_MIXED
    B = np.zeros(nf)
    with pytest.raises(_lib.IllumError) as e:  # host buffers
        _lib.call("illum_radiosity_solve_noalloc", Fm.h, api._ptr(E), 1, m, 1e-8, 50,
                  api._ptr(B), api._ptr(resid), api.C.byref(it), 0)
    assert e.value.code == _lib.ILLUM_ERR_ARG
    with pytest.raises(_lib.IllumError) as e:  # max_iter above the bound
        _lib.call("illum_radiosity_solve_noalloc", Fm.h, api._ptr(Ed), 1, m, 1e-8, 300,
                  api._ptr(Bd), api._ptr(resid), api.C.byref(it), dev)
    assert e.value.code == _lib.ILLUM_ERR_ARG
    # 1,280 facets x 2,048 right-hand sides needs more than 64 MiB
    S = 2048
    Eb = api.DeviceArray(ctx, (S, nf))
    Bb = api.DeviceArray(ctx, (S, nf))
    with pytest.raises(_lib.IllumError) as e:
        _lib.call("illum_radiosity_solve_noalloc", Fm.h, api._ptr(Eb), S, m, 1e-8, 50,
                  api._ptr(Bb), api._ptr(resid), api.C.byref(it), dev)
    assert e.value.code == _lib.ILLUM_ERR_CAPACITY
    # and the next call on the same GPU still works
    a = _solve(api, "illum_radiosity_solve", Fm, Ed, Bd, 1, "mixed", 1e-10, 100)
    Bn = api.DeviceArray(ctx, (nf,))
    b = _solve(api, "illum_radiosity_solve_noalloc", Fm, Ed, Bn, 1, "mixed", 1e-10, 100)
    assert a[2] == b[2] and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
