"""Seeded synthetic workloads of the five measured configurations, shared by
bench.py and the tests so both run exactly the same inputs.

  C1  5,120 facets (icosphere subdivided 4 times), 40 craters; horizons at 64
      azimuth intervals (nphi 65); 1 sun; albedo 0.1; Gauss-Seidel to 1e-10;
      Crank-Nicolson 1 rotation of 256 steps, 40 layers to 0.5 m, thermal
      inertia 200, emissivity 0.95
  C2  81,920 facets (subdivided 6 times), 150 craters; full form-factor
      assembly; horizons at 360 intervals (nphi 361)
  C3  the C2 mesh, 512 sun directions uniform on the +z hemisphere, per-facet
      albedo U(0.05, 0.15), radiosity to 1e-8 relative residual
  C4  327,680 facets (subdivided 7 times), 150 craters; 4,096 CN steps over 4
      rotations; per-facet layer counts U{32..96}; thermal inertia lognormal
      (median 200, sigma 0.5)
  C5  64 bodies, triaxial ellipsoids with axis ratios U(0.6, 1.0), subdivided
      3 to 7 times (1,280 to 327,680 facets, log-uniform), 32 suns each, one
      rotation of 512 CN steps

Every ellipsoid has semi-axes (13, 11, 9) km unless stated, craters of depth
U(0.05, 0.3) km and width U(0.2, 1.5) km.  The sun sits at 227,390,024 km
(illum_cli.cpp:337); the rotation period is Vesta's, 5.342 h (runjob.sh:3-4);
volumetric heat capacity 1.19177e6 J m^-3 K^-1 (share/config.yaml:21).
"""
import numpy as np

from . import api

SUN_KM = 227390024.0
ROT_PERIOD_S = 5.342 * 3600
RHOC = 1.19177e6
TI = 200.0
EMISS = 0.95
AU_KM = 149.5978707e6

MESHES = {"C1": (4, 40), "C2": (6, 150), "C3": (6, 150), "C4": (7, 150)}


def mesh(config, seed=0):
    """(V, F) of a configuration's seeded cratered ellipsoid."""
    subdiv, ncraters = MESHES[config]
    return api.synth_ellipsoid(subdiv, ncraters=ncraters, seed=seed)


def facet_albedo(nf, seed=0):
    """Per-facet albedo U(0.05, 0.15)."""
    return np.random.default_rng(seed).uniform(0.05, 0.15, nf)


def layer_grid(nz, depth=0.5):
    """nz exponentially spaced layer depths (m) down to `depth`."""
    return depth * (np.exp(np.linspace(0, 3, nz)) - 1 + 0.01) / (np.exp(3) - 1 + 0.01)


def hemisphere_suns(S, seed):
    """3 x S sun positions (km), directions uniform on the +z hemisphere."""
    rng = np.random.default_rng(seed)
    z = rng.uniform(0, 1, S)
    ph = rng.uniform(0, 2 * np.pi, S)
    r = np.sqrt(1 - z * z)
    return SUN_KM * np.stack([r * np.cos(ph), r * np.sin(ph), z])


def orbit_sun(k, steps_per_rot, inclination_deg=20.0):
    """Sun position (km) at step k of a rotation: the body spins about z under
    a sun 20 degrees above its equator."""
    a = 2 * np.pi * k / steps_per_rot
    inc = np.deg2rad(inclination_deg)
    return SUN_KM * np.array([np.cos(a) * np.cos(inc), np.sin(a) * np.cos(inc), np.sin(inc)])


def c4_layers(nf, seed=0):
    """C4's ragged layering: (grids, grid_of_facet, ti_scale) with one grid per
    layer count 32..96, facets drawn uniformly over them, and a lognormal
    thermal-inertia factor (median 1, sigma 0.5) on the base inertia."""
    rng = np.random.default_rng(seed)
    nzs = np.arange(32, 97)
    grids = [layer_grid(n) for n in nzs]
    gid = rng.integers(0, len(nzs), nf).astype(np.int32)
    tis = np.exp(rng.normal(0, 0.5, nf))
    return grids, gid, tis


def insolation(normals, k0, nsteps, steps_per_rot, albedo=0.1):
    """Absorbed flux Q (W/m^2), nsteps x nf, at steps k0+1 .. k0+nsteps of the
    orbit_sun rotation, without shadowing: the load of the conduction-only
    runs."""
    S0 = 1372 * (AU_KM / SUN_KM) ** 2 * (1 - albedo)
    inc = np.deg2rad(20.0)
    ang = 2 * np.pi * np.arange(k0 + 1, k0 + nsteps + 1) / steps_per_rot
    s = np.stack([np.cos(ang) * np.cos(inc), np.sin(ang) * np.cos(inc), np.full(nsteps, np.sin(inc))])
    return np.ascontiguousarray(S0 * np.clip(s.T @ np.asarray(normals, np.float64).T, 0, None))


def c5_suns(nbody, S=32, seed=0):
    """3 x (S nbody) sun positions (km): body b's S suns (columns b S ..
    b S + S - 1) are S evenly spaced steps of one rotation under a sun at a
    per-body elevation U(0, 40) degrees above the equator."""
    rng = np.random.default_rng(seed + 7)
    inc = rng.uniform(0.0, 40.0, nbody)
    return np.concatenate([np.stack([orbit_sun(k, S, inc[b]) for k in range(S)], 1)
                           for b in range(nbody)], 1)


def c5_bodies(nbody=64, seed=0, max_subdiv=7):
    """C5's uneven batch: [(V, F)] of nbody seeded ellipsoids, axis ratios
    U(0.6, 1.0) on 13 km, subdivision levels uniform on 3..max_subdiv (facet
    counts log-uniform), 10 + 20 (subdiv - 3) craters each."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(nbody):
        sub = int(rng.integers(3, max_subdiv + 1))
        ax = tuple(13.0 * rng.uniform(0.6, 1.0, 3))
        out.append(api.synth_ellipsoid(sub, axes=ax, ncraters=10 + 20 * (sub - 3),
                                       seed=1000 * seed + b))
    return out
