"""Builders shared by the tests: the same system on the device (through the
reference-style API) and in the CPU oracle."""

from __future__ import annotations

import numpy as np

from oracle import oracle as O

ISO = dict(kn=2.0e5, kt=2.0 / 7.0 * 2.0e5, gamman=50.0, gammat=25.0, xmu=0.5)


def pair_for(tm, dom, style, **kw):
    c = dict(ISO)
    c.update(kw)
    if style == "hooke/history":
        return tm.GranHookeHistory(dom, c["kn"], c["gamman"], c["kt"], c["gammat"], c["xmu"])
    if style == "hertz/history":
        return tm.GranHertzHistory(dom, c["kn"], c["gamman"], c["kt"], c["gammat"], c["xmu"])
    if style == "granular":
        return tm.PairGranular(dom, normal="hertz/material", normal_coeffs=[1e7, 0.5, 0.3],
                               tangential="mindlin", tangential_coeffs=[None, 1.0, 0.5],
                               damping="viscoelastic")
    raise ValueError(style)


def oracle_pair_for(style, **kw):
    c = dict(ISO)
    c.update(kw)
    if style in ("hooke/history", "hertz/history"):
        return O.legacy_pair(style, kn=c["kn"], gamman=c["gamman"], kt=c["kt"],
                             gammat=c["gammat"], xmu=c["xmu"])
    return O.granular_pair("hertz/material", [1e7, 0.5, 0.3], "mindlin", [None, 1.0, 0.5],
                           "viscoelastic")


def device_sim(w, style="hooke/history", precision="f64", servo=True, skin=0.1, dt=1e-4,
               max_nb=64, cundall=None, servo_kw=None, check=True, every=1, **pair_kw):
    import taichimps_amd as tm
    n = len(w["x"])
    dom = tm.Domain(w["boxlo"], w["boxhi"], ("p", "p", "p"))
    atom = tm.AtomSystem(max_atoms=n, float_type=precision)
    atom.add_particles(x=w["x"], radius=w["radius"], density=w["density"], v=w["v"],
                       omega=w["omega"], tag=w["tag"])
    nl = tm.NeighborList(dom, skin=skin, max_atoms=n, max_neighbors_per_atom=max_nb)
    nl.check = check
    nl.every = every
    hist = tm.ContactHistory(n, max_nb)
    sim = tm.Simulation(dom, atom, nl, hist, pair_for(tm, dom, style, **pair_kw), dt=dt)
    sim.add_fix(tm.FixNVESphere(dom))
    if cundall is not None:
        sim.add_fix(tm.FixDampingCundall(dom, *cundall))
    if servo:
        sim.add_fix(tm.FixDeformPressure(dom, **(servo_kw or dict(p_target=1.0, p_gain=1e-4))))
    return sim


def oracle_sim(w, style="hooke/history", servo=True, skin=0.1, max_nb=64, cundall=None,
               servo_kw=None, check=True, every=1, **pair_kw):
    o = O.from_workload(w, max_nb=max_nb)
    o.set_neigh(skin, 0, every, check)
    o.set_pair(oracle_pair_for(style, **pair_kw))
    if cundall is not None:
        o.set_cundall(*cundall)
    if servo:
        o.set_servo(O.servo_params(**(servo_kw or dict(p_target=1.0, p_gain=1e-4))))
    return o


def device_state(sim):
    n = sim.atom.nlocal
    a = sim.atom
    sim.domain.pull()
    return {
        "x": a.x.to_numpy()[:n], "v": a.v.to_numpy()[:n], "omega": a.omega.to_numpy()[:n],
        "f": a.f.to_numpy()[:n], "torque": a.torque.to_numpy()[:n],
        "virial": a.virial.to_numpy()[:n],
        "boxlo": np.array(sim.domain.boxlo), "boxhi": np.array(sim.domain.boxhi),
    }


def oracle_state(o):
    s = o.atoms()
    lo, hi, _ = o.box()
    s["boxlo"], s["boxhi"] = lo, hi
    return s


def max_rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)
