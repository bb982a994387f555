"""Writes tests/golden/cn_golden.npz: conductionQ outputs of the reference's own
Fortran (conductionQ.f + tridag.for, compiled into oracle/_ref by
`make -C oracle ref`) for seeded inputs.  Run in the build container:

    python tests/golden/make_cn_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402


def cases():
    rng = np.random.default_rng(20181201)
    out = []
    # share/config.yaml:5-21 parameter set (24 layers, ti 70, rhoc 1.19177e6, T0 233)
    cfg_z = np.array([
        0.00103331794606226, 0.00227329948133698, 0.00376127732366664, 0.00554685073446223,
        0.00768953882741695, 0.0102607645389626, 0.0133462353928174, 0.0170488004174431,
        0.021491878446994, 0.0268235720824551, 0.0332216044450084, 0.0408992432800723,
        0.050112409882149, 0.0611682098046411, 0.0744351697116316, 0.0903555216000201,
        0.109459943866086, 0.132385250585366, 0.159895618648501, 0.192908060324264,
        0.232522990335179, 0.280060906348277, 0.337106405563995, 0.405561004622856])
    out.append(dict(name="config_yaml", z=cfg_z, ti=np.full(24, 70.0), rhoc=np.full(24, 1.19177e6),
                    dt=600.0, emiss=0.999, T0=233.0, Fgeo=0.0, nsteps=64,
                    Q=300 * np.maximum(0, np.sin(np.linspace(0, 4 * np.pi, 64)))))
    # 40 exponentially spaced layers to 0.5 m, TI 200, emissivity 0.95, 256 steps
    z40 = 0.5 * (np.exp(np.linspace(0, 3, 40)) - 1 + 0.01) / (np.exp(3) - 1 + 0.01)
    out.append(dict(name="exp40", z=z40, ti=np.full(40, 200.0), rhoc=np.full(40, 1.19177e6),
                    dt=5.342 * 3600 / 256, emiss=0.95, T0=233.0, Fgeo=0.0, nsteps=256,
                    Q=400 * np.maximum(0, np.cos(np.linspace(0, 2 * np.pi, 256)))))
    # irregular grid, layered properties, geothermal flux
    for k in range(3):
        nz = int(rng.integers(5, 60))
        z = np.cumsum(rng.uniform(0.001, 0.02, nz))
        out.append(dict(name=f"random{k}", z=z, ti=rng.uniform(30, 400, nz),
                        rhoc=rng.uniform(0.8e6, 1.6e6, nz), dt=float(rng.uniform(30, 900)),
                        emiss=float(rng.uniform(0.8, 1.0)), T0=float(rng.uniform(150, 300)),
                        Fgeo=float(rng.uniform(0, 0.05)), nsteps=32,
                        Q=rng.uniform(0, 600, 32)))
    return out


def main():
    if not O.ref_available():
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    arrays = {}
    names = []
    for c in cases():
        nz = len(c["z"])
        T = np.full(nz, c["T0"])
        Ts = c["T0"]
        Qprev = 0.0
        Tr, Tsr, Fsr = [], [], []
        for q in c["Q"]:
            T, Ts, Fs = O.ref_conductionQ(nz, c["z"], c["dt"], Qprev, q, T, c["ti"], c["rhoc"],
                                          c["emiss"], Ts, c["Fgeo"])
            Qprev = q
            Tr.append(T.copy())
            Tsr.append(Ts)
            Fsr.append(Fs)
        n = c["name"]
        names.append(n)
        for key in ("z", "ti", "rhoc", "Q"):
            arrays[f"{n}.{key}"] = np.asarray(c[key], np.float64)
        arrays[f"{n}.scalars"] = np.array([c["dt"], c["emiss"], c["T0"], c["Fgeo"]])
        arrays[f"{n}.T"] = np.array(Tr)
        arrays[f"{n}.Tsurf"] = np.array(Tsr)
        arrays[f"{n}.Fsurf"] = np.array(Fsr)
    arrays["names"] = np.array(names)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cn_golden.npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, {k: v.shape for k, v in arrays.items() if k.endswith(".T")})


if __name__ == "__main__":
    main()
