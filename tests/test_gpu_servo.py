"""The fix deform/pressure style table on the device against the oracle
(reference src/taichimps/EXTRA_FIX/deform_pressure.py:582-702, the
`remap v` partner-velocity shift domain.py:388-414 and the PBC velocity
correction :502-571), FP64, bit for bit.

Every style code the device servo implements runs in a whole device-resident
Simulation.run on a dense packing (contacts everywhere, so the pressure
styles see a contact virial), and through the per-stage path
(Simulation.step). The oracle pins of the same styles are in
tests/test_oracle_servo.py.
"""

import numpy as np
import pytest

import simhelpers as H
from oracle import oracle as O
from taichimps_amd import workloads

pytestmark = pytest.mark.gpu

P = dict(style="pressure", ptarget=0.4, pgain=2e-3)

# name -> (device FixDeformPressure kwargs, oracle servo_params kwargs)
CASES = {
    "trate": dict(axes={d: {"style": "trate", "rate": -0.8} for d in "xyz"}),
    "vel_none": dict(axes={"x": {"style": "vel", "vel": -0.6}, "y": {"style": "vel", "vel": 0.3},
                           "z": {"style": "none"}}),
    "wiggle_pressure": dict(axes={"x": {"style": "wiggle", "amplitude": 0.04, "tperiod": 4e-3},
                                  "y": {"style": "wiggle", "amplitude": -0.02, "tperiod": 3e-3},
                                  "z": dict(P)}),
    "final_delta_scale": dict(axes={"x": {"style": "final", "flo": 0.03, "fhi": 8.75},
                                    "y": {"style": "delta", "dlo": 0.02, "dhi": -0.05},
                                    "z": {"style": "scale", "scale": 0.99}}),
    "couple_xy": dict(axes={"x": dict(P), "y": dict(P), "z": {"style": "erate", "rate": -0.5}},
                      couple="xy"),
    "couple_yz": dict(axes={"x": {"style": "none"}, "y": dict(P), "z": dict(P)}, couple="yz"),
    "couple_xz": dict(axes={"x": dict(P), "y": {"style": "trate", "rate": 0.3}, "z": dict(P)},
                      couple="xz"),
    "pressure_mean": dict(axes={"x": {"style": "pressure/mean", "ptarget": 0.4, "pgain": 2e-3},
                                "y": dict(P), "z": {"style": "none"}}),
    "max_rate": dict(axes={d: dict(P, pgain=1.0) for d in "xyz"}, max_rate=0.05),
    "normalize": dict(axes={"x": dict(P), "y": {"style": "pressure", "ptarget": 0.0, "pgain": 2e-3},
                            "z": dict(P)}, max_rate=0.2, normalize=True),
    "remap_none": dict(axes={d: {"style": "erate", "rate": -2.0} for d in "xyz"}, remap="none"),
    "remap_v": dict(axes={d: {"style": "erate", "rate": -4.0} for d in "xyz"}, remap="v"),
}


def _device_kw(spec):
    kw = dict(spec)
    if "normalize" in kw:
        kw["normalize_pressure"] = kw.pop("normalize")
    return kw


def _oracle_kw(spec):
    return dict(spec)


def _workload():
    """A dense 8^3 packing (spacing 1.1 with radii up to 0.55, plus speeds of
    order 1) shifted so the top layer sits 0.06 below the upper faces: a
    shrinking box pushes it across them, and the rebuild's wrap (with the
    `remap v` velocity correction) is exercised."""
    w = workloads.lattice(8, 8, 8, 1.1, 31, "uniform")
    rng = np.random.default_rng(32)
    w["v"] = rng.normal(0.0, 1.0, size=w["v"].shape)
    w["x"] = w["x"] + (0.55 - 0.06)
    return w


@pytest.mark.parametrize("name", sorted(CASES))
def test_servo_style_run_bitexact(name):
    spec = CASES[name]
    w = _workload()
    steps = 120
    sim = H.device_sim(w, style="hertz/history", servo=True, skin=0.01,
                       servo_kw=_device_kw(spec))
    sim.run(steps)
    assert sim.last_run_device_resident
    o = H.oracle_sim(w, style="hertz/history", servo=False, skin=0.01)
    o.set_servo(O.servo_params(**_oracle_kw(spec)))
    assert o.run(steps, 1e-4) == 0
    dev, orc = H.device_state(sim), H.oracle_state(o)
    for k in ("x", "v", "omega", "f", "torque", "virial", "boxlo", "boxhi"):
        assert np.array_equal(dev[k], orc[k]), (name, k, H.max_rel(dev[k], orc[k]))
    lo0, hi0 = np.asarray(w["boxlo"]), np.asarray(w["boxhi"])
    assert not np.array_equal(dev["boxhi"] - dev["boxlo"], hi0 - lo0), name  # the box moved
    assert sim.neighbor.build_count == o.counters()["build_count"] > 1
    _, _, hr = o.box()
    sim.domain.pull()
    assert np.array_equal(np.asarray(sim.domain.h_rate)[:3], hr[:3]), name


@pytest.mark.parametrize("name", ["remap_v", "final_delta_scale", "couple_xz", "normalize"])
def test_servo_style_stage_path_bitexact(name):
    """The same styles through Simulation.step (one stage launch each), as the
    reference's unit tests drive them."""
    spec = CASES[name]
    w = _workload()
    steps = 40
    sim = H.device_sim(w, style="hertz/history", servo=True, skin=0.01,
                       servo_kw=_device_kw(spec))
    sim.setup_run(steps)
    sim.init_simulation()
    for _ in range(steps):
        sim.step()
    o = H.oracle_sim(w, style="hertz/history", servo=False, skin=0.01)
    o.set_servo(O.servo_params(**_oracle_kw(spec)))
    assert o.run(steps, 1e-4) == 0
    dev, orc = H.device_state(sim), H.oracle_state(o)
    for k in ("x", "v", "omega", "f", "boxlo", "boxhi"):
        assert np.array_equal(dev[k], orc[k]), (name, k, H.max_rel(dev[k], orc[k]))


def test_remap_v_wraps_atoms_and_corrects_their_velocity():
    """Under remap v with a shrinking box some atoms leave through the upper
    faces and are wrapped at a rebuild; their velocities then carry the
    image's lattice velocity (the oracle's closed-form pin is in
    tests/test_oracle_servo.py)."""
    w = _workload()
    spec = CASES["remap_v"]
    sim = H.device_sim(w, style="hertz/history", servo=True, skin=0.01,
                       servo_kw=_device_kw(spec))
    sim.run(120)
    x = sim.atom.x.to_numpy()[:len(w["x"])]
    sim.domain.pull()
    lo = np.asarray(sim.domain.boxlo)
    # atoms that started at the top layer now sit near the lower faces
    top = w["x"][:, 0] > w["boxhi"][0] - 0.1
    assert top.any() and (x[top, 0] < lo[0] + 0.5).any()
