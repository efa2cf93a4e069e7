"""Frozen synthetic inputs for the DEM step benchmarks and parity cases.

The contract (one generator, fixed draw order) is:

    rng = numpy.random.default_rng(seed)
    1. jitter  U(-0.04, 0.04), shape (N, 3)
    2. radii   (per case)
    3. v       N(0, 0.1), shape (N, 3)

omega = 0, density 1 (rmass = 4/3 pi r^3, computed by AtomSystem.add_particles
exactly as the reference does, atom.py:110-111), tags 1..N in lattice order
(x slowest, z fastest), sites (i + 0.5) * a, box [0, n_d * a] per axis, fully
periodic.

Cases:
  isotropic_hooke_8k   20^3 lattice, a = 1.1, radii U[0.45, 0.55], seed 1001
  isotropic_hertz_1m   100^3 lattice, a = 1.1, radii U[0.45, 0.55], seed 1002
  triaxial_4m          160 x 125 x 200, a = 1.3, log-normal radii (median 0.5,
                       sigma 0.15, clipped [0.35, 0.7]), seed 1003
  adjoint_32k          32^3, a = 1.1, radii U[0.45, 0.55], seed 1004

Ensembles (ensemble()): 256 members from SeedSequence(1005).spawn(256); member
k draws, from default_rng(child k), a target count N log-uniform in
[2048, 65536], kn U[1e5, 1e6], xmu U[0.2, 0.7], target pressure U[0.5, 2.0],
and a lattice seed; its lattice is nx x ny x nz with nx = ny = round(N^(1/3))
and nz = round(N / nx^2) (a = 1.1, radii U[0.45, 0.55]), and even members use
gran/hooke/history, odd ones gran/hertz/history.
"""

from __future__ import annotations

import math

import numpy as np


def lattice(nx: int, ny: int, nz: int, a: float, seed: int, radii: str = "uniform",
            count: int | None = None) -> dict:
    """Jittered simple-cubic lattice; the first ``count`` sites when given."""
    n_full = nx * ny * nz
    n = n_full if count is None else int(count)
    rng = np.random.default_rng(seed)
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    sites = np.stack([(ix.ravel() + 0.5) * a, (iy.ravel() + 0.5) * a,
                      (iz.ravel() + 0.5) * a], axis=1)[:n]
    jitter = rng.uniform(-0.04, 0.04, size=(n, 3))
    if radii == "uniform":
        radius = rng.uniform(0.45, 0.55, size=n)
    elif radii == "lognormal":
        radius = np.clip(rng.lognormal(mean=math.log(0.5), sigma=0.15, size=n), 0.35, 0.7)
    else:
        raise ValueError(f"unknown radii distribution {radii!r}")
    v = rng.normal(0.0, 0.1, size=(n, 3))
    return {
        "x": sites + jitter,
        "v": v,
        "omega": np.zeros((n, 3)),
        "radius": radius,
        "density": 1.0,
        "tag": np.arange(1, n + 1, dtype=np.int32),
        "boxlo": np.zeros(3),
        "boxhi": np.array([nx * a, ny * a, nz * a]),
    }


# Contact and run settings shared by the isotropic cases (kt = 2/7 kn and
# gammat = gamman / 2 are the NULL shorthands of input.py:444-472).
ISOTROPIC = dict(kn=2.0e5, kt=2.0 / 7.0 * 2.0e5, gamman=50.0, gammat=25.0, xmu=0.5,
                 dt=1e-4, skin=0.1, p_target=1.0, p_gain=1e-4)

CASES = {
    "isotropic_hooke_8k": dict(dims=(20, 20, 20), a=1.1, seed=1001, radii="uniform",
                               style="hooke/history", steps=1000),
    "isotropic_hertz_1m": dict(dims=(100, 100, 100), a=1.1, seed=1002, radii="uniform",
                               style="hertz/history", steps=200),
    "triaxial_4m": dict(dims=(160, 125, 200), a=1.3, seed=1003, radii="lognormal",
                        style="granular", steps=1000),
    "adjoint_32k": dict(dims=(32, 32, 32), a=1.1, seed=1004, radii="uniform",
                        style="hertz/history", steps=2000),
}


def case(name: str, dims: tuple[int, int, int] | None = None) -> dict:
    """The named case's atoms, optionally on a smaller lattice of the same recipe."""
    c = CASES[name]
    nx, ny, nz = dims if dims is not None else c["dims"]
    w = lattice(nx, ny, nz, c["a"], c["seed"], c["radii"])
    w["name"] = name
    w["style"] = c["style"]
    w["steps"] = c["steps"]
    return w


def ensemble(count: int = 256, seed: int = 1005, n_min: int = 2048, n_max: int = 65536) -> list:
    """The ensemble members' atoms and settings (see the module docstring).
    Each entry is a lattice() dict plus style, kn, xmu and p_target."""
    out = []
    for k, child in enumerate(np.random.SeedSequence(seed).spawn(count)):
        rng = np.random.default_rng(child)
        target = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        kn = float(rng.uniform(1.0e5, 1.0e6))
        xmu = float(rng.uniform(0.2, 0.7))
        p_target = float(rng.uniform(0.5, 2.0))
        lat_seed = int(rng.integers(0, 2**31 - 1))
        nx = max(2, int(round(target ** (1.0 / 3.0))))
        nz = max(2, int(round(target / (nx * nx))))
        w = lattice(nx, nx, nz, 1.1, lat_seed, "uniform")
        w.update(name=f"ensemble_{k}", style="hooke/history" if k % 2 == 0 else "hertz/history",
                 kn=kn, xmu=xmu, p_target=p_target)
        out.append(w)
    return out
