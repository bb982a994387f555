"""illum_cli (C++ driver over the C ABI): tasks and output files of the
reference driver (illum_cli.cpp:270-510), checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "illum_amd", "illum_cli")


def run(args, tmp_path):
    r = subprocess.run([CLI, *args, "--output_dir", str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def load_sparse(path):
    with open(path, "rb") as f:
        assert f.readline() == b"ARMA_SPM_BIN_FN008\n"
        r, c, nnz = map(int, f.readline().split())
        val = np.frombuffer(f.read(8 * nnz), np.float64)
        ri = np.frombuffer(f.read(8 * nnz), np.uint64)
        cp = np.frombuffer(f.read(8 * (c + 1)), np.uint64)
    return r, c, cp.astype(np.int64), ri.astype(np.int32), val


def _synth(subdiv, ncr, seed):
    from illum_amd import api
    return api.synth_ellipsoid(subdiv, ncraters=ncr, seed=seed)


def test_cli_form_factors_file(tmp_path):
    out = run(["form_factors", "--synth", "3,40,2", "--albedo", "0.12"], tmp_path)
    assert "nnz" in out
    n, c, cp, ri, v = load_sparse(tmp_path / "form_factors.bin")
    V, F = _synth(3, 40, 2)
    g = O.geometry(V, F)
    bits, _ = O.visibility(g)
    ocp, ori, ov = O.form_factors(g, bits, albedo=0.12)
    assert n == c == len(F)
    assert np.array_equal(cp, ocp) and np.array_equal(ri, ori)
    np.testing.assert_allclose(v, ov, rtol=1e-12)


def test_cli_horizons_and_radiosity_thermal(tmp_path):
    from illum_amd.armaio import load_mat
    run(["horizons", "--synth", "3,40,1", "--nphi", "33"], tmp_path)
    hz = load_mat(tmp_path / "horizons.bin")
    V, F = _synth(3, 40, 1)
    g = O.geometry(V, F)
    ho, _ = O.horizons(g, nphi=33)
    assert np.array_equal(hz, ho)
    sun = tmp_path / "sun.txt"
    sun.write_text("1.5e11 2.0e10 1.0e11\n1.2e11 -8.0e10 1.1e11\n")
    depths = ",".join(str(x) for x in np.linspace(0.01, 0.4, 20))
    out = run(["radiosity", "--synth", "3,40,1", "-r", "-t", "--horizon_file", str(tmp_path / "horizons.bin"),
               "--sun_pos_file", str(sun), "--depths", depths, "--ti", "70", "--T0", "233",
               "--dt", "600", "--albedo", "0.3", "--print_residual"], tmp_path)
    assert "rel res" in out
    rad = load_mat(tmp_path / "rad.bin").ravel()
    therm = load_mat(tmp_path / "therm.bin").ravel()
    # oracle replay of the reference loop (illum_cli.cpp:403-456) with GS to 1e-13
    disk = O.fib_spiral(100)
    bits, _ = O.visibility(g)
    cp, ri, v = O.form_factors(g, bits, albedo=0.3)
    z = np.linspace(0.01, 0.4, 20)
    nf = len(F)
    T = np.full((21, nf), 233.0)
    Qprev = np.zeros(nf)
    Fs = np.zeros(nf)
    for s in (np.array([1.5e8, 2.0e7, 1.0e8]), np.array([1.2e8, -8.0e7, 1.1e8])):
        E = -Fs + O.direct(g, ho, s, disk)
        B, _, _ = O.gauss_seidel(nf, cp, ri, v, E, max_sweeps=200, rel_tol=1e-13)
        T, Qprev, Fs, _ = O.thermal_run([0, 20], z, np.full(20, 70.0), np.full(20, 1.19177e6), T,
                                        Qprev, B[None, :], 600.0, 0.999)
    np.testing.assert_allclose(rad, B, rtol=1e-8, atol=1e-8 * np.abs(B).max())
    # B agrees to the solve tolerance; CN propagates that to ~1e-5 K at 700 K
    np.testing.assert_allclose(therm, T[0], rtol=0, atol=1e-4)
