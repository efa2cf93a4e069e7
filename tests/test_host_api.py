"""Host-side logic of the reference-style API, without a GPU: coefficient
derivations against the oracle's independent restatement, argument order,
configuration errors raised as the reference raises them, and the workload
generator's reproducibility.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O
import taichimps_amd as tm
from taichimps_amd import workloads
from taichimps_amd.EXTRA_FIX.deform_pressure import FixDeformPressure


def _dom():
    return tm.Domain([0.0] * 3, [10.0] * 3, ("p", "p", "p"))


def _same(a, b):
    assert bytes(a) == bytes(b), (
        [(f, getattr(a, f), getattr(b, f)) for f, _ in a._fields_ if getattr(a, f) != getattr(b, f)])


@pytest.mark.parametrize("style,cls", [("hooke/history", tm.GranHookeHistory),
                                       ("hertz/history", tm.GranHertzHistory)])
@pytest.mark.parametrize("dampflag", [0, 1])
def test_legacy_pair_params_match_oracle(style, cls, dampflag):
    """Constructor order kn, gamman, kt, gammat, xmu; dampflag 0 zeroes gammat."""
    p = cls(_dom(), 2e5, 50.0, 2e5 * 2 / 7, 25.0, 0.5, dampflag=dampflag).params()
    q = O.legacy_pair(style, kn=2e5, gamman=50.0, kt=2e5 * 2 / 7, gammat=25.0, xmu=0.5,
                      dampflag=dampflag)
    _same(p, q)
    assert (p.gammat == 0.0) == (dampflag == 0)


GRANULAR_CASES = [
    ("hooke", [2e5, 30.0], "none", None, "viscoelastic"),
    ("hertz", [2e5, 0.5], "linear_history", [5e4, 1.0, 0.4], "velocity"),
    ("hertz/material", [1e7, 0.8, 0.3], "mindlin", [None, 1.0, 0.5], "viscoelastic"),
    ("hertz/material", [71.6e9, 0.95, 0.23], "mindlin", [None, 1.0, 0.3], "coeff_restitution"),
    ("hooke", [1e5, 0.9], "linear_nohistory", [1.0, 0.3], "coeff_restitution"),
    ("hertz", [1e5, 0.7], "linear_history", [2e4, 1.0, 0.6], "tsuji"),
    ("hooke", [1e5, 3.0], "none", None, "mass_velocity"),
]


@pytest.mark.parametrize("normal,nc,tang,tc,damp", GRANULAR_CASES)
def test_granular_params_match_oracle(normal, nc, tang, tc, damp):
    p = tm.PairGranular(_dom(), normal, nc, tangential=tang, tangential_coeffs=tc,
                        damping=damp).params()
    q = O.granular_pair(normal, nc, tangential=tang, tangential_coeffs=tc, damping=damp)
    _same(p, q)


def test_hertz_material_stiffness_closed_form():
    """kn = 4/3 E* with E* = E / (2 (1 - nu^2)) for like materials; mindlin
    NULL kt = 8 G* with G* = E / (4 (2 - nu)(1 + nu))."""
    E, nu = 1e7, 0.3
    p = tm.PairGranular(_dom(), "hertz/material", [E, 0.5, nu], tangential="mindlin",
                        tangential_coeffs=[None, 1.0, 0.5]).params()
    np.testing.assert_allclose(p.g_kn, 4.0 / 3.0 * E / (2 * (1 - nu * nu)), rtol=1e-14)
    np.testing.assert_allclose(p.g_kt, 8.0 * E / (4 * (2 - nu) * (1 + nu)), rtol=1e-14)


def test_coeff_restitution_hooke_damping_closed_form():
    """Linear spring: restitution e gives damp = -2 ln e / sqrt(pi^2 + ln^2 e)."""
    e = 0.8
    p = tm.PairGranular(_dom(), "hooke", [1e5, e], damping="coeff_restitution").params()
    np.testing.assert_allclose(p.g_damp, -2 * math.log(e) / math.hypot(math.pi, math.log(e)),
                               rtol=1e-15)


@pytest.mark.parametrize("kwargs,exc", [
    (dict(normal="dmt", normal_coeffs=[1, 2]), ValueError),
    (dict(normal="hooke", normal_coeffs=[1e5, 0.5], damping="tsuji"), None),
    (dict(normal="hooke", normal_coeffs=[1e5, 1.5], damping="tsuji"), ValueError),
    (dict(normal="hooke", normal_coeffs=[1e5, 0.5], tangential="mindlin",
          tangential_coeffs=[None, 1.0, 0.5]), ValueError),
    (dict(normal="hooke", normal_coeffs=[1e5]), ValueError),
])
def test_granular_configuration_errors(kwargs, exc):
    if exc is None:
        tm.PairGranular(_dom(), **kwargs)
    else:
        with pytest.raises(exc):
            tm.PairGranular(_dom(), **kwargs)


def test_deform_pressure_legacy_isotropic_form():
    """deform_pressure.py:190-199: all axes pressure, gain 1e-4, couple xyz."""
    f = FixDeformPressure(_dom(), p_target=1.0)
    p = f.params()
    assert f.couple == "xyz" and p.couple == 1
    assert list(p.style) == [8, 8, 8]
    assert list(p.pgain) == [1e-4] * 3 and list(p.ptarget) == [1.0] * 3
    _same(p, O.servo_params(p_target=1.0, p_gain=1e-4))


@pytest.mark.parametrize("kwargs,exc", [
    (dict(axes={"x": {"style": "pressure", "ptarget": 1.0, "pgain": 0.0}}), ValueError),
    (dict(axes={"x": {"style": "erate/rescale", "rate": 1.0}}), ValueError),
    (dict(axes={"x": {"style": "erate", "rate": 1.0}}, couple="xy"), ValueError),
    (dict(axes={"x": {"style": "pressure", "ptarget": 1.0, "pgain": 1e-4},
                "y": {"style": "pressure", "ptarget": 2.0, "pgain": 1e-4}}, couple="xy"),
     ValueError),
    (dict(axes={"x": {"style": "pressure", "ptarget": 1.0, "pgain": 1e-4}}, max_rate=-1.0),
     ValueError),
    (dict(axes={"xy": {"style": "erate", "rate": 1.0}}), NotImplementedError),
    (dict(axes={"x": {"style": "volume"}}), NotImplementedError),
    (dict(axes={"x": {"style": "pressure", "ptarget": 1.0, "pgain": 1e-4}}, nevery=0),
     ValueError),
])
def test_deform_pressure_configuration_errors(kwargs, exc):
    with pytest.raises(exc):
        FixDeformPressure(_dom(), **kwargs)


def test_deform_pressure_per_axis_params_match_oracle():
    axes = {"x": {"style": "pressure", "ptarget": 1.0, "pgain": 1e-4},
            "y": {"style": "pressure", "ptarget": 1.0, "pgain": 1e-4},
            "z": {"style": "erate", "rate": -1e-3}}
    p = FixDeformPressure(_dom(), axes=axes, couple="xy", max_rate=0.5).params()
    q = O.servo_params(axes=axes, couple="xy", max_rate=0.5)
    _same(p, q)


def test_domain_tilt_is_rejected():
    with pytest.raises(NotImplementedError):
        tm.Domain([0.0] * 3, [1.0] * 3, ("p", "p", "p"), tilt=[0.1, 0.0, 0.0])


def test_workload_generator_is_seeded_and_matches_spec():
    """Draw order jitter, radii, velocities; sites (i + 0.5) a; box [0, n a]."""
    w1 = workloads.case("isotropic_hooke_8k")
    w2 = workloads.case("isotropic_hooke_8k")
    for k in ("x", "v", "radius"):
        assert np.array_equal(w1[k], w2[k])
    assert len(w1["x"]) == 8000
    np.testing.assert_allclose(w1["boxhi"], [22.0] * 3)
    rng = np.random.default_rng(1001)
    jitter = rng.uniform(-0.04, 0.04, size=(8000, 3))
    radii = rng.uniform(0.45, 0.55, size=8000)
    idx = np.arange(8000)
    site = np.stack([idx // 400, (idx // 20) % 20, idx % 20], axis=1)
    np.testing.assert_allclose(w1["x"], (site + 0.5) * 1.1 + jitter, rtol=0, atol=1e-12)
    assert np.array_equal(w1["radius"], radii)
    assert np.array_equal(w1["tag"], np.arange(1, 8001))


def test_ensemble_workload_is_seeded_and_matches_spec():
    """256 members from SeedSequence(1005).spawn(256): N_k log-uniform in
    [2048, 65536], the first N_k sites of a ceil-sized lattice, kn, xmu and
    target P in their ranges, hooke and hertz alternating."""
    e1 = workloads.ensemble()
    e2 = workloads.ensemble()
    assert len(e1) == 256
    for a, b in zip(e1, e2):
        assert np.array_equal(a["x"], b["x"]) and a["kn"] == b["kn"]
    ns = np.array([len(w["x"]) for w in e1])
    assert ns.min() >= 2048 and ns.max() <= 65536
    # log-uniform: about half of the members below the geometric midpoint
    assert 0.35 < np.mean(ns < np.sqrt(2048 * 65536)) < 0.65
    assert all(1e5 <= w["kn"] <= 1e6 and 0.2 <= w["xmu"] <= 0.7
               and 0.5 <= w["p_target"] <= 2.0 for w in e1)
    assert [w["style"] for w in e1[:4]] == ["hooke/history", "hertz/history"] * 2
    # member k's draws come from its own spawned generator
    rng = np.random.default_rng(np.random.SeedSequence(1005).spawn(256)[3])
    target = np.exp(rng.uniform(np.log(2048), np.log(65536)))
    assert e1[3]["kn"] == rng.uniform(1.0e5, 1.0e6)
    n_k = int(round(target))
    assert len(e1[3]["x"]) == n_k
    nx = int(np.ceil(n_k ** (1.0 / 3.0)))
    nz = int(np.ceil(n_k / (nx * nx)))
    assert nx * nx * (nz - 1) < n_k <= nx * nx * nz
    np.testing.assert_allclose(e1[3]["boxhi"], [nx * 1.1, nx * 1.1, nz * 1.1])
    # the first N_k sites in lattice order (x slowest, z fastest)
    full = workloads.lattice(nx, nx, nz, 1.1, 0, "uniform")
    site = np.round(e1[3]["x"] / 1.1 - 0.5)
    assert np.array_equal(site, np.round(full["x"][:n_k] / 1.1 - 0.5))
