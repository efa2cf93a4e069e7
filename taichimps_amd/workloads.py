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
k draws, from default_rng(child k), a count N_k = round(exp U(ln 2048,
ln 65536)), kn U[1e5, 1e6], xmu U[0.2, 0.7], target pressure U[0.5, 2.0],
and a lattice seed; its atoms are the first N_k sites of an nx x ny x nz
lattice with nx = ny = ceil(N_k^(1/3)) and nz = ceil(N_k / nx^2) (a = 1.1,
radii U[0.45, 0.55], the box the whole lattice's), and even members use
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


def _icbrt_ceil(n: int) -> int:
    """ceil(n^(1/3)) in integers (a float cube root of a perfect cube can land
    an ulp above the integer)."""
    c = int(round(n ** (1.0 / 3.0)))
    while c * c * c < n:
        c += 1
    while c > 1 and (c - 1) ** 3 >= n:
        c -= 1
    return c


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
        n_k = int(round(target))
        nx = max(2, _icbrt_ceil(n_k))
        nz = max(2, -(-n_k // (nx * nx)))
        w = lattice(nx, nx, nz, 1.1, lat_seed, "uniform", count=n_k)
        w.update(name=f"ensemble_{k}", style="hooke/history" if k % 2 == 0 else "hertz/history",
                 kn=kn, xmu=xmu, p_target=p_target)
        out.append(w)
    return out


# ---------------------------------------------------------------- systems
# The benchmark configurations as reference-API systems (Simulation with its
# fixes), shared by bench.py and the parity tests.

def build_isotropic(w: dict, precision: str = "f64", style: str = "hertz/history"):
    """Isotropic compression: gran/hertz/history (or hooke) with the ISOTROPIC
    coefficients, nve/sphere and deform/pressure P = 1.0 (gain 1e-4, couple
    xyz, remap x), 64 list slots per atom."""
    import taichimps_amd as tm
    c = ISOTROPIC
    n = len(w["x"])
    dom = tm.Domain(w["boxlo"], w["boxhi"], ("p", "p", "p"))
    atom = tm.AtomSystem(max_atoms=n, float_type=precision)
    atom.add_particles(x=w["x"], radius=w["radius"], density=w["density"], v=w["v"],
                       omega=w["omega"], tag=w["tag"])
    nl = tm.NeighborList(dom, skin=c["skin"], max_atoms=n, max_neighbors_per_atom=64)
    hist = tm.ContactHistory(n, 64)
    cls = tm.GranHookeHistory if style == "hooke/history" else tm.GranHertzHistory
    pair = cls(dom, c["kn"], c["gamman"], c["kt"], c["gammat"], c["xmu"])
    sim = tm.Simulation(dom, atom, nl, hist, pair, dt=c["dt"])
    sim.add_fix(tm.FixNVESphere(dom))
    sim.add_fix(tm.FixDeformPressure(dom, p_target=c["p_target"], p_gain=c["p_gain"]))
    return sim


# The triaxial test's settings (the configuration leaves them open; these are
# the ones its benchmark records).
TRIAXIAL = dict(E=1.0e7, damp=0.5, poisson=0.3, xt=1.0, mu=0.5, cundall=0.1, dt=1e-4,
                skin=0.1, p_target=1.0, p_gain=1e-4, z_erate=-1e-3)


def build_triaxial(w: dict, precision: str = "f64"):
    """The triaxial test: PairGranular hertz/material E = 1e7, damp 0.5,
    nu 0.3, viscoelastic damping, mindlin (NULL, 1.0, 0.5); damping/cundall
    0.1 0.1; x and y pressure servo P = 1.0 (gain 1e-4, couple none), z erate
    -1e-3; dt 1e-4, skin 0.1."""
    import taichimps_amd as tm
    c = TRIAXIAL
    n = len(w["x"])
    dom = tm.Domain(w["boxlo"], w["boxhi"], ("p", "p", "p"))
    atom = tm.AtomSystem(max_atoms=n, float_type=precision)
    atom.add_particles(x=w["x"], radius=w["radius"], density=w["density"], v=w["v"],
                       omega=w["omega"], tag=w["tag"])
    nl = tm.NeighborList(dom, skin=c["skin"], max_atoms=n, max_neighbors_per_atom=64)
    hist = tm.ContactHistory(n, 64)
    pair = tm.PairGranular(dom, normal="hertz/material",
                           normal_coeffs=[c["E"], c["damp"], c["poisson"]],
                           tangential="mindlin", tangential_coeffs=[None, c["xt"], c["mu"]],
                           damping="viscoelastic")
    sim = tm.Simulation(dom, atom, nl, hist, pair, dt=c["dt"])
    sim.add_fix(tm.FixNVESphere(dom))
    sim.add_fix(tm.FixDampingCundall(dom, c["cundall"], c["cundall"]))
    sim.add_fix(tm.FixDeformPressure(dom, axes={
        "x": {"style": "pressure", "ptarget": c["p_target"], "pgain": c["p_gain"]},
        "y": {"style": "pressure", "ptarget": c["p_target"], "pgain": c["p_gain"]},
        "z": {"style": "erate", "rate": c["z_erate"]}}))
    return sim


def build_member(w: dict, precision: str = "f32split"):
    """An ensemble member (ensemble()): its lattice, pair style and
    coefficients (kt = 2/7 kn, gamman 50, gammat 25), nve/sphere and
    deform/pressure at its own target."""
    import taichimps_amd as tm
    c = ISOTROPIC
    n = len(w["x"])
    dom = tm.Domain(w["boxlo"], w["boxhi"], ("p", "p", "p"))
    atom = tm.AtomSystem(max_atoms=n, float_type=precision)
    atom.add_particles(x=w["x"], radius=w["radius"], density=w["density"], v=w["v"],
                       omega=w["omega"], tag=w["tag"])
    nl = tm.NeighborList(dom, skin=c["skin"], max_atoms=n, max_neighbors_per_atom=64)
    hist = tm.ContactHistory(n, 64)
    cls = tm.GranHookeHistory if w["style"] == "hooke/history" else tm.GranHertzHistory
    pair = cls(dom, w["kn"], c["gamman"], 2.0 / 7.0 * w["kn"], c["gammat"], w["xmu"])
    sim = tm.Simulation(dom, atom, nl, hist, pair, dt=c["dt"])
    sim.add_fix(tm.FixNVESphere(dom))
    sim.add_fix(tm.FixDeformPressure(dom, p_target=w["p_target"], p_gain=c["p_gain"]))
    return sim
