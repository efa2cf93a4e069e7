"""The triaxial and ensemble configurations against the FP64 oracle, at
reduced size (GPU).

Triaxial: the 4M-sphere recipe (log-normal radii clipped to [0.35, 0.7] on a
1.3 lattice, seed 1003) on a 24 x 19 x 30 lattice (13,680 spheres) with the
benchmark's contact law and fixes: PairGranular hertz/material + viscoelastic
+ mindlin, damping/cundall 0.1, x/y pressure servo, z erate. FP64 is bitwise
equal to the oracle; FP32-split is held to the FP64 oracle with the
tolerances stated in its test.

Ensemble: members of the real 256-member generator (workloads.ensemble, seed
1005) advanced together as an Ensemble in FP32-split; each member's pressure,
positions and box are compared with that member's own FP64 oracle run.
"""

import numpy as np
import pytest

import simhelpers as H
from oracle import oracle as O
from taichimps_amd import workloads

pytestmark = pytest.mark.gpu

TRIAX_DIMS = (24, 19, 30)


def _box_err(dev, orc):
    """Largest difference of a face position, relative to the box length."""
    length = orc["boxhi"] - orc["boxlo"]
    return float(max(np.abs(dev["boxlo"] - orc["boxlo"]).max(),
                     np.abs(dev["boxhi"] - orc["boxhi"]).max()) / length.min())


def triaxial_oracle(w):
    c = workloads.TRIAXIAL
    o = O.from_workload(w)
    o.set_neigh(c["skin"])
    o.set_pair(O.granular_pair("hertz/material", [c["E"], c["damp"], c["poisson"]], "mindlin",
                               [None, c["xt"], c["mu"]], "viscoelastic"))
    o.set_cundall(c["cundall"], c["cundall"])
    o.set_servo(O.servo_params(axes={
        "x": {"style": "pressure", "ptarget": c["p_target"], "pgain": c["p_gain"]},
        "y": {"style": "pressure", "ptarget": c["p_target"], "pgain": c["p_gain"]},
        "z": {"style": "erate", "rate": c["z_erate"]}}, couple="none"))
    return o


def test_triaxial_reduced_fp64_bitwise():
    w = workloads.case("triaxial_4m", dims=TRIAX_DIMS)
    assert w["radius"].max() > 0.6 and w["radius"].min() < 0.4  # polydisperse
    steps = 150
    sim = workloads.build_triaxial(w, "f64")
    sim.run(steps)
    o = triaxial_oracle(w)
    assert o.run(steps, workloads.TRIAXIAL["dt"]) == 0
    dev, orc = H.device_state(sim), H.oracle_state(o)
    for k in ("x", "v", "omega", "f", "torque", "virial", "boxlo", "boxhi"):
        assert np.array_equal(dev[k], orc[k]), (k, H.max_rel(dev[k], orc[k]))
    assert np.array_equal(sim.atom._native.last_pressure(), o.last_pressure_tensor())
    assert sim.neighbor.build_count == o.counters()["build_count"] > 1


def test_triaxial_reduced_fp32_split_vs_fp64_oracle():
    """FP32 contacts and integration with FP64 virial and servo, 150 steps
    of the overlapping polydisperse start: per-axis pressure within 1e-4
    relative, positions within 1e-4 absolute (the lattice spacing is 1.3),
    forces within 2e-2 of the largest force, faces within 1e-7 of the box
    length."""
    w = workloads.case("triaxial_4m", dims=TRIAX_DIMS)
    steps = 150
    sim = workloads.build_triaxial(w, "f32split")
    sim.run(steps)
    o = triaxial_oracle(w)
    assert o.run(steps, workloads.TRIAXIAL["dt"]) == 0
    dev, orc = H.device_state(sim), H.oracle_state(o)
    p, pref = sim.atom._native.last_pressure(), o.last_pressure_tensor()
    perr = np.abs(p[:3] - pref[:3]) / np.abs(pref[:3])
    xerr = np.abs(dev["x"] - orc["x"]).max()
    ferr = np.abs(dev["f"] - orc["f"]).max() / np.abs(orc["f"]).max()
    berr = _box_err(dev, orc)
    print(f"triaxial split: pressure rel {perr}, x abs {xerr:.3e}, f rel {ferr:.3e}, "
          f"box rel {berr:.3e}")
    assert perr.max() < 1e-4, perr
    assert xerr < 1e-4
    assert ferr < 2e-2
    assert berr < 1e-7


def _member_oracle(w):
    c = workloads.ISOTROPIC
    o = O.from_workload(w)
    o.set_neigh(c["skin"])
    o.set_pair(O.legacy_pair(w["style"], kn=w["kn"], gamman=c["gamman"], kt=2.0 / 7.0 * w["kn"],
                             gammat=c["gammat"], xmu=w["xmu"]))
    o.set_servo(O.servo_params(p_target=w["p_target"], p_gain=c["p_gain"]))
    return o


def _member_floor(w, steps, seeds=(1, 2, 3, 4), amp=1e-6):
    """How far FP64 runs of the member drift from its unperturbed FP64 run
    when the input positions carry U(+-amp) noise: (mean-pressure relative,
    position absolute), worst over the seeds. Over 100 steps the pressure
    responds linearly to it (1e-7 and 1e-6 give drifts ten times apart), at
    about 20 per unit of position; FP32-split's own position error over
    those steps is 2e-6 (the offsets' rounding, accumulated by the drift)."""
    o = _member_oracle(w)
    assert o.run(steps, workloads.ISOTROPIC["dt"]) == 0
    p0, x0 = o.last_pressure_tensor()[:3].mean(), o.atoms()["x"]
    dp = dx = 0.0
    for seed in seeds:
        wp = dict(w)
        rng = np.random.default_rng(seed)
        wp["x"] = w["x"] + rng.uniform(-amp, amp, size=w["x"].shape)
        o = _member_oracle(wp)
        assert o.run(steps, workloads.ISOTROPIC["dt"]) == 0
        dp = max(dp, abs(o.last_pressure_tensor()[:3].mean() - p0) / abs(p0))
        dx = max(dx, float(np.abs(o.atoms()["x"] - x0).max()))
    return dp, dx


def test_ensemble_members_vs_fp64_oracle():
    """Eight members of the 256-member ensemble (the six smallest and the two
    largest, both pair styles), 100 FP32-split steps in one Ensemble run.
    Each member against its own FP64 oracle run: mean pressure within 1e-5
    relative and positions within 2e-5 absolute, or within twice the drift of
    the member's FP64 run from inputs perturbed by U(+-1e-6) (_member_floor)
    where that is larger: a 2,000-atom packing's pressure moves by ~2e-5 for
    a 1e-6 shift of its positions, about what FP32 offsets amount to; faces
    within 1e-7 of the box length."""
    from taichimps_amd.ensemble import Ensemble
    ws = workloads.ensemble()
    order = np.argsort([len(w["x"]) for w in ws])
    pick = [int(k) for k in list(order[:6]) + list(order[-2:])]
    sel = [ws[k] for k in pick]
    assert {w["style"] for w in sel} == {"hooke/history", "hertz/history"}
    assert min(len(w["x"]) for w in sel) >= 2048 and max(len(w["x"]) for w in sel) <= 65536
    steps = 100
    sims = [workloads.build_member(w, "f32split") for w in sel]
    Ensemble(sims).run(steps)
    failures = []
    for k, (w, sim) in enumerate(zip(sel, sims)):
        o = _member_oracle(w)
        assert o.run(steps, workloads.ISOTROPIC["dt"]) == 0
        dev, orc = H.device_state(sim), H.oracle_state(o)
        p, pref = sim.atom._native.last_pressure(), o.last_pressure_tensor()
        perr = abs(p[:3].mean() - pref[:3].mean()) / abs(pref[:3].mean())
        xerr = float(np.abs(dev["x"] - orc["x"]).max())
        berr = _box_err(dev, orc)
        fp, fx = _member_floor(w, steps)
        print(f"member {pick[k]} ({w['style']}, N {len(w['x'])}): mean P rel {perr:.3e} "
              f"(floor {fp:.3e}), x abs {xerr:.3e} (floor {fx:.3e}), box rel {berr:.3e}")
        assert sim.timestep == steps
        if perr > max(1e-5, 2.0 * fp) or xerr > max(2e-5, 2.0 * fx) or berr > 1e-7:
            failures.append((pick[k], perr, fp, xerr, fx, berr))
    assert not failures, failures
