"""The tangent-linear restatement (oracle/tlm.cpp), the collision-proof
checker of the device adjoint: CPU pins.

Its FP64 form must be the oracle's run bit for bit (positions, velocities,
pressure tensor, number of builds), on runs with neighbor rebuilds; its
tangent form must carry those same values and derivatives that agree with
central finite differences of the oracle at the prescribed 1e-6 relative step,
on runs short and small enough for finite differences to converge.
"""

from __future__ import annotations

import numpy as np
import pytest

import simhelpers as H
from oracle import oracle as O
from taichimps_amd import workloads

C = workloads.ISOTROPIC
COEFFS = np.array([C["kn"], C["kt"], C["gamman"], C["gammat"], C["xmu"]])
MEAN_P = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0]) / 3.0


def _case(seed=31, n=8, speed=2.0):
    w = workloads.lattice(n, n, n, 1.0, seed, "uniform")
    rng = np.random.default_rng(seed + 1)
    w["v"] = rng.normal(0.0, speed, size=w["v"].shape)
    w["omega"] = rng.normal(0.0, 3.0, size=w["v"].shape)
    return w


@pytest.mark.parametrize("style", ["hertz/history", "hooke/history"])
def test_fp64_form_is_the_oracle_run_bitwise(style):
    w = _case()
    o = H.oracle_sim(w, style=style, skin=0.02, servo_kw=dict(p_target=1.0, p_gain=2e-4))
    assert o.run(80, 1e-4) == 0
    x, v, t, builds = O.tlm_forward(w, COEFFS, 80, style=style, skin=0.02, p_gain=2e-4)
    a = o.atoms()
    assert np.array_equal(x, a["x"]) and np.array_equal(v, a["v"])
    assert np.array_equal(t, o.last_pressure_tensor())
    assert builds == o.counters()["build_count"] > 3
    t2, _, builds2 = O.tlm_tangent(w, COEFFS, 80, style=style, skin=0.02, p_gain=2e-4)
    assert np.array_equal(t2, t) and builds2 == builds


def _oracle_pmean(w, coeffs, radius, style, steps):
    w2 = dict(w)
    w2["radius"] = radius
    o = O.from_workload(w2)
    o.set_neigh(0.02)
    kn, kt, gn, gt, xmu = coeffs
    o.set_pair(O.legacy_pair(style, kn=kn, gamman=gn, kt=kt, gammat=gt, xmu=xmu))
    o.set_servo(O.servo_params(p_target=1.0, p_gain=2e-4))
    assert o.run(steps, 1e-4) == 0
    return float(MEAN_P @ o.last_pressure_tensor())


def _persistent_case(seed=61, n=6, a=0.95):
    """A compressed lattice whose contacts neither open nor close over the
    run: hooke's normal force jumps at contact onset, so finite differences
    across an onset are no test of a derivative."""
    w = workloads.lattice(n, n, n, a, seed, "uniform")
    rng = np.random.default_rng(seed + 1)
    ix, iy, iz = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    sites = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1) * a + 0.5 * a
    w["x"] = sites + rng.uniform(-0.005, 0.005, size=sites.shape)
    w["radius"] = rng.uniform(0.49, 0.51, size=len(sites))
    w["v"] = rng.normal(0.0, 0.01, size=sites.shape)
    w["omega"] = rng.normal(0.0, 2.0, size=sites.shape)
    return w


@pytest.mark.parametrize("style,xmu", [("hertz/history", 0.5), ("hooke/history", 0.05),
                                       ("hooke/history", 50.0)])
def test_tangent_matches_central_differences_of_the_oracle(style, xmu):
    """60 steps with rebuilds: d<P>/d(kn, kt, gamman, gammat, xmu) and the
    derivative along a random radius direction (at fixed density) within
    1e-5 of central differences at relative step 1e-6; hertz on a colliding
    packing (its force is continuous at onset), hooke on one whose contacts
    persist, sliding (xmu 0.05) or sticking (xmu 50): a contact that crosses
    the Coulomb limit inside a difference is a kink, no test of a
    derivative."""
    w = _case(seed=41, n=6) if style == "hertz/history" else _persistent_case()
    coeffs = COEFFS.copy()
    coeffs[4] = xmu
    steps = 60
    radius = np.asarray(w["radius"], dtype=np.float64)
    u = np.random.default_rng(7).normal(size=len(radius))
    _, d, builds = O.tlm_tangent(w, coeffs, steps, dradius=u, style=style, skin=0.02,
                                 p_gain=2e-4)
    assert builds > 1
    tl = MEAN_P @ d
    fd = np.zeros(5)
    for k in range(5):
        h = 1e-6 * coeffs[k]
        cp, cm = coeffs.copy(), coeffs.copy()
        cp[k] += h
        cm[k] -= h
        fd[k] = (_oracle_pmean(w, cp, radius, style, steps)
                 - _oracle_pmean(w, cm, radius, style, steps)) / (2.0 * h)
    # the log-scaled sensitivities theta_k dP/dtheta_k: the vector to 1e-5 of
    # its largest entry, each entry at least 1e-3 of the largest to 1e-5 of
    # itself (a difference resolves a smaller one only to the run's rounding,
    # which the dynamics amplify)
    s_tl, s_fd = coeffs * tl[:5], coeffs * fd
    big = np.abs(s_fd).max()
    assert np.abs(s_tl - s_fd).max() <= 1e-5 * big, (s_tl, s_fd)
    for k in range(5):
        if abs(s_fd[k]) >= 1e-3 * big:
            assert abs(tl[k] - fd[k]) <= 1e-5 * abs(fd[k]), (k, tl[k], fd[k])
    eps = 1e-6 * np.abs(radius).max()
    fd = (_oracle_pmean(w, coeffs, radius + eps * u, style, steps)
          - _oracle_pmean(w, coeffs, radius - eps * u, style, steps)) / (2.0 * eps)
    assert abs(tl[5] - fd) <= 1e-5 * abs(fd), (tl[5], fd)
