"""Ensembles (taichimps_amd/ensemble.py, csrc/batch.hip): every member's state
after Ensemble.run is bitwise the state its own Simulation.run gives it.

Members differ in size (none a multiple of the block size), box, pair style
(hooke/history, hertz/history, PairGranular), contact coefficients, servo
target and servo presence; a small skin makes each member rebuild its list at
its own steps, so the members' redo paths and ragged ends are exercised.
"""

import numpy as np
import pytest

import simhelpers as H
from taichimps_amd import workloads
from taichimps_amd.ensemble import Ensemble

pytestmark = pytest.mark.gpu

# (lattice side, spacing, seed, style, kn, xmu, p_target or None, skin)
MEMBERS = [
    (7, 1.0, 101, "hooke/history", 3.0e5, 0.3, 1.5, 0.02),
    (9, 1.02, 102, "hertz/history", 8.0e5, 0.6, 0.7, 0.02),
    (5, 1.0, 103, "hertz/history", 1.2e5, 0.25, None, 0.03),
    (11, 1.05, 104, "hooke/history", 6.0e5, 0.5, 2.0, 0.1),
    (6, 1.0, 105, "granular", 2.0e5, 0.5, 1.0, 0.02),
    (8, 0.98, 106, "hertz/history", 2.0e5, 0.45, 1.2, 0.015),
]


def member(spec):
    side, a, seed, style, kn, xmu, p, skin = spec
    w = workloads.lattice(side, side, side, a, seed, "uniform")
    rng = np.random.default_rng(seed + 7)
    w["v"] = rng.normal(0.0, 0.5, size=w["v"].shape)
    servo_kw = None if p is None else dict(p_target=p, p_gain=1e-4)
    kw = {} if style == "granular" else dict(kn=kn, kt=2.0 / 7.0 * kn, xmu=xmu)
    return H.device_sim(w, style=style, precision="f32split", servo=p is not None, skin=skin,
                        servo_kw=servo_kw, **kw)


def members():
    return [member(spec) for spec in MEMBERS]


def assert_same(a, b, what):
    sa, sb = H.device_state(a), H.device_state(b)
    for k in ("x", "v", "omega", "f", "torque", "virial", "boxlo", "boxhi"):
        assert np.array_equal(sa[k], sb[k]), f"{what}: {k} differs"
    assert a.timestep == b.timestep, what
    assert a.neighbor.build_count == b.neighbor.build_count, what
    ca, cb = a.atom._native.counters(), b.atom._native.counters()
    for k in ("ntouch", "nslots", "last_build_step", "servo_step_count"):
        assert ca[k] == cb[k], f"{what}: {k}"


@pytest.mark.parametrize("autotune", [False, True])
def test_members_match_their_single_runs(autotune):
    ens_sims, solo = members(), members()
    ens = Ensemble(ens_sims)
    ens.run(90, autotune=autotune)
    for k, s in enumerate(solo):
        s.run(90)
    rebuilt = [s.neighbor.build_count for s in solo]
    assert max(rebuilt) > 2 and len(set(rebuilt)) > 1, rebuilt
    for k in range(len(solo)):
        assert_same(ens_sims[k], solo[k], f"member {k} ({MEMBERS[k][3]})")
    # a continued run (no Verlet::setup) matches a continued single run
    for s in ens_sims + solo:
        s.run_pre = False
    ens.run(37)
    for s in solo:
        s.run(37)
    for k in range(len(solo)):
        assert_same(ens_sims[k], solo[k], f"member {k} continued")
    t = ens.tuning()
    assert t["redone_steps"] >= 0
    if t["redone_steps"] > 0:
        assert t["redo_passes"] >= 1 and t["redo_ms"] > 0.0, t
    if autotune:
        rates = t["rates_atom_steps_per_s"]
        assert all(v > 0 for v in rates.values()), rates
        best = max(rates, key=rates.get)
        assert best == f"{t['segment_length']}/{t['block_order']}"


def test_uneven_sizes_and_a_single_member():
    """One member alone, then members whose runs end at different steps of a
    segment (different rebuild counts), over a run shorter than a segment."""
    sims, solo = [member(MEMBERS[1])], [member(MEMBERS[1])]
    Ensemble(sims).run(5)
    solo[0].run(5)
    assert_same(sims[0], solo[0], "single member")
    sims, solo = members()[:3], members()[:3]
    Ensemble(sims).run(3)
    for s in solo:
        s.run(3)
    for k in range(3):
        assert_same(sims[k], solo[k], f"short run member {k}")


def test_zero_steps_is_setup_only():
    sims, solo = members()[:2], members()[:2]
    Ensemble(sims).run(0)
    for s in solo:
        s.run(0)
    for k in range(2):
        assert_same(sims[k], solo[k], f"setup member {k}")


def test_long_rows_and_forced_rebuilds():
    """Members whose list rows exceed the single-pass scratch of 32 (a skin of
    1.2 spacings: about 45 partners per atom), rebuilt every third step
    without the skin check, beside a member with it: the redone steps take
    the two-pass list fill, and several members redo in the same pass."""
    def sims():
        out = []
        for side, seed, style in ((8, 201, "hooke/history"), (9, 202, "hertz/history")):
            w = workloads.lattice(side, side, side, 1.0, seed, "uniform")
            w["v"] = np.random.default_rng(seed).normal(0.0, 0.5, size=w["v"].shape)
            out.append(H.device_sim(w, style=style, precision="f32split", servo=True, skin=1.2,
                                    check=False, every=3, kn=4.0e5, kt=2.0 / 7.0 * 4.0e5,
                                    xmu=0.5))
        out.append(member(MEMBERS[0]))
        return out

    ens_sims, solo = sims(), sims()
    Ensemble(ens_sims).run(20)
    for s in solo:
        s.run(20)
    assert solo[0].neighbor.build_count >= 6
    counts = solo[1].atom._native.counters()
    assert counts["nslots"] > 32 * solo[1].atom.nlocal // 2
    for k in range(len(solo)):
        assert_same(ens_sims[k], solo[k], f"member {k}")


def test_rejects_what_it_does_not_run():
    w = workloads.lattice(4, 4, 4, 1.0, 5, "uniform")
    fp64 = H.device_sim(w, precision="f64")
    with pytest.raises(NotImplementedError):
        Ensemble([fp64])
    damped = H.device_sim(w, precision="f32split", cundall=(0.1, 0.1))
    with pytest.raises(NotImplementedError):
        Ensemble([damped])
    a = H.device_sim(w, precision="f32split")
    with pytest.raises(ValueError):
        Ensemble([a, a])
    b = H.device_sim(w, precision="f32split", dt=2e-4)
    with pytest.raises(ValueError):
        Ensemble([H.device_sim(w, precision="f32split"), b])
    with pytest.raises(ValueError):
        Ensemble([])


def test_member_gradients_do_not_depend_on_the_thread_count():
    """ensemble.gradients drives members' adjoints on host threads at once;
    each member's pressure and gradients are bitwise those of a one-thread
    pass."""
    from taichimps_amd import adjoint as ADJ
    from taichimps_amd.ensemble import gradients
    ws = [workloads.lattice(side, side, side, 1.05, 200 + side, "uniform") for side in (6, 9, 7, 8)]
    builders = [ADJ.isotropic_builder(w, style="hertz/history" if k % 2 else "hooke/history",
                                      skin=0.02) for k, w in enumerate(ws)]
    coeffs = [np.array([2e5 * (1 + 0.1 * k), 2.0 / 7.0 * 2e5, 50.0, 25.0, 0.4]) for k in range(4)]
    radii = [w["radius"] for w in ws]
    one = gradients(builders, coeffs, radii, 40, 15, workers=1)
    many = gradients(builders, coeffs, radii, 40, 15, workers=4)
    for (t1, c1, r1), (t2, c2, r2) in zip(one, many):
        assert np.array_equal(t1, t2) and np.array_equal(c1, c2) and np.array_equal(r1, r2)
        assert np.all(np.isfinite(c1)) and np.abs(c1).max() > 0.0
