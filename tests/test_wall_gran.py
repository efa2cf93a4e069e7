"""fix wall/gran (reference: EXTRA_FIX/wall_gran.py).

CPU tests pin the oracle's wall contact by closed forms of one sphere against
a wall: the hooke normal force with mass-velocity damping, the tangential
history at stick, the Coulomb cap with the history rescale at slip, the
history-free viscous force, the moment arm of the full radius, the wall's
shear and wiggle motion, and the history reset on separation. GPU tests
compare the HIP kernel with the oracle bit for bit, per call and over whole
runs through Simulation (the wall puts the run on the stage path).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O

KN, KT, GN, GT, MU = 1.0e5, 2.0e4, 10.0, 5.0, 0.5
DT = 1.0e-4


def _one(x, v, omega=(0.0, 0.0, 0.0), r=0.5):
    o = O.OracleSim(1)
    rr = np.array([r])
    o.set_atoms(np.array([x], dtype=float), np.array([v], dtype=float),
                np.array([omega], dtype=float), rr, O.mass_from_radius(rr, 1.0))
    o.set_box((0.0, 0.0, -1.0), (2.0, 2.0, 3.0), (1, 1, 0))
    o.set_neigh(0.1)
    return o, O.mass_from_radius(rr, 1.0)[0]


def _force(o, idx, hu=True):
    o.clear_forces()
    o.wall_post_force(idx, DT, hu)
    a = o.atoms()
    return a["f"][0], a["torque"][0]


def test_normal_force_and_stick_closed_form_low_wall():
    o, m = _one((1.0, 1.0, 0.45), (0.1, 0.0, -0.2))
    w = o.add_wall(2, -1, 0.0, KN, GN, KT, GT, MU)
    f, t = _force(o, w)
    fn = KN * 0.05 + m * GN * 0.2
    fs = -KT * 0.1 * DT - m * GT * 0.1
    np.testing.assert_allclose(f, [fs, 0.0, fn], rtol=1e-13)
    # moment arm r (not r - delta/2): torque = -r n x fs with n = +z
    np.testing.assert_allclose(t, [0.0, -0.5 * fs, 0.0], rtol=1e-13)
    sh, tc = o.wall_history(w)
    np.testing.assert_allclose(sh[0], [0.1 * DT, 0.0, 0.0], rtol=1e-13)
    assert tc[0] == 1


def test_high_wall_normal_points_down_and_history_projects_out_the_normal():
    o, m = _one((1.0, 1.55, 1.0), (0.0, 0.3, 0.0), omega=(0.0, 0.0, 2.0))
    w = o.add_wall(1, 1, 2.0, KN, GN, KT, GT, MU)
    f, t = _force(o, w)
    delta = 0.5 - (2.0 - 1.55)
    fn = KN * delta - m * GN * (0.3 * -1.0)
    # n = -y; vr.n = -0.3; vt = 0; vtr = -(r omega x n) = -(0.5*2 z x -y) = -(1.0 x) ... in x
    vtr = -np.cross(0.5 * np.array([0.0, 0.0, 2.0]), [0.0, -1.0, 0.0])
    fs = -KT * (vtr * DT) - m * GT * vtr
    np.testing.assert_allclose(f, [fs[0], -fn, 0.0], rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(t, -0.5 * np.cross([0.0, -1.0, 0.0], fs), rtol=1e-13)
    # the spin is opposed by the wall's friction torque
    assert t[2] * 2.0 < 0.0


def test_slip_caps_at_coulomb_and_rescales_history():
    o, m = _one((1.0, 1.0, 0.4999), (5.0, 0.0, 0.0))
    w = o.add_wall(2, -1, 0.0, KN, GN, KT, GT, MU)
    f, _ = _force(o, w)
    fn = KN * (0.5 - 0.4999)
    assert f[2] == pytest.approx(fn, rel=1e-13)
    # uncapped |fs| would be kt |v| dt + m gt |v| = 23.1 > mu fn = 5: slip
    assert abs(np.hypot(f[0], f[1]) - MU * fn) <= 1e-12 * MU * fn
    sh, _ = o.wall_history(w)
    fdamp = -m * GT * np.array([5.0, 0.0, 0.0])
    np.testing.assert_allclose(sh[0], -(np.array([f[0], f[1], 0.0]) - fdamp) / KT, rtol=1e-12)


def test_history_free_style_is_viscous_capped():
    o, m = _one((1.0, 1.0, 0.45), (0.01, 0.0, 0.0))
    w = o.add_wall(2, -1, 0.0, KN, GN, KT, GT, MU, history=False)
    f, _ = _force(o, w)
    assert f[0] == pytest.approx(-m * GT * 0.01, rel=1e-13)
    # m gt |v| = 131 above mu fn = 5: capped at the Coulomb limit
    o2, _ = _one((1.0, 1.0, 0.4999), (50.0, 0.0, 0.0))
    w2 = o2.add_wall(2, -1, 0.0, KN, GN, KT, GT, MU, history=False)
    f2, _ = _force(o2, w2)
    assert f2[0] == pytest.approx(-MU * KN * (0.5 - 0.4999), rel=1e-13)


def test_separation_resets_history_and_setup_does_not_accumulate():
    o, _ = _one((1.0, 1.0, 0.45), (0.1, 0.0, 0.0))
    w = o.add_wall(2, -1, 0.0, KN, GN, KT, GT, MU)
    _force(o, w, hu=False)
    sh, tc = o.wall_history(w)
    assert tc[0] == 1 and np.array_equal(sh[0], np.zeros(3))
    _force(o, w)
    assert o.wall_history(w)[0][0][0] > 0.0
    o.set_atoms(np.array([[1.0, 1.0, 0.6]]), np.zeros((1, 3)), np.zeros((1, 3)),
                np.array([0.5]), O.mass_from_radius(np.array([0.5]), 1.0))
    f, _ = _force(o, w)
    sh, tc = o.wall_history(w)
    assert tc[0] == 0 and np.array_equal(sh[0], np.zeros(3)) and np.array_equal(f, np.zeros(3))


def test_sheared_wall_drags_a_resting_sphere():
    o, m = _one((1.0, 1.0, 0.45), (0.0, 0.0, 0.0))
    w = o.add_wall(2, -1, 0.0, KN, GN, KT, GT, MU, shear_axis=0, shear_vel=0.2)
    f, _ = _force(o, w)
    assert f[0] == pytest.approx(KT * 0.2 * DT + m * GT * 0.2, rel=1e-13)


def test_wiggle_moves_the_wall_by_the_closed_form():
    amp, per = 0.01, 20 * DT
    o, m = _one((1.0, 1.0, 0.45), (0.0, 0.0, 0.0))
    w = o.add_wall(2, -1, 0.0, KN, GN, KT, GT, MU, wiggle_amplitude=amp, wiggle_period=per)
    f0, _ = _force(o, w)
    assert f0[2] == pytest.approx(KN * 0.05, rel=1e-13)  # elapsed 0: wall at 0, at rest
    f1, _ = _force(o, w)  # elapsed dt
    om = 2.0 * math.pi / per
    coord = amp * (1.0 - math.cos(om * DT))
    vn = amp * om * math.sin(om * DT)
    assert f1[2] == pytest.approx(KN * (0.5 - (0.45 - coord)) - m * GN * (0.0 - vn), rel=1e-12)


# ------------------------------------------------------------------ GPU

def _cases():
    out = []
    for axis in range(3):
        for side in (-1, 1):
            out.append(dict(axis=axis, side=side, history=True))
    out += [dict(axis=2, side=-1, history=False), dict(axis=0, side=1, history=True,
                                                      limit_damping=True),
            dict(axis=2, side=-1, history=True, shear_axis=1, shear_vel=0.3),
            dict(axis=1, side=-1, history=True, wiggle_amplitude=0.02, wiggle_period=0.003)]
    return out


def _box_system(seed, n=400):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.2, 3.8, size=(n, 3))
    r = rng.uniform(0.25, 0.35, size=n)
    v = rng.normal(0.0, 0.5, size=(n, 3))
    w = rng.normal(0.0, 3.0, size=(n, 3))
    return x, r, v, w


@pytest.mark.gpu
@pytest.mark.parametrize("case", _cases())
def test_wall_kernel_bitexact_per_call(case):
    """Five post_force calls on spheres spread through the wall's band, with
    random velocities and spins, history carried between calls."""
    import taichimps_amd as tm
    x, r, v, om = _box_system(7)
    n = len(x)
    coord = 0.25 if case["side"] == -1 else 3.75
    dom = tm.Domain((0.0, 0.0, 0.0), (4.0, 4.0, 4.0), ("f", "f", "f"))
    atom = tm.AtomSystem(max_atoms=n)
    atom.add_particles(x=x, radius=r, density=1.0, v=v, omega=om)
    atom._bind_domain(dom)
    kw = {k: case[k] for k in ("shear_axis", "shear_vel", "wiggle_amplitude",
                              "wiggle_period", "limit_damping") if k in case}
    fix = tm.FixWallGran(dom, case["axis"], case["side"], coord, KN, GN, KT, GT, MU,
                         history=case["history"], max_atoms=n, **kw)
    o = O.OracleSim(n)
    o.set_atoms(x, v, om, r, O.mass_from_radius(r, 1.0))
    o.set_box((0.0,) * 3, (4.0,) * 3, (0, 0, 0))
    wi = o.add_wall(case["axis"], case["side"], coord, KN, GN, KT, GT, MU,
                    history=case["history"], **kw)
    for call in range(5):
        atom.clear_forces()
        o.clear_forces()
        fix.history_update = call != 0
        fix.post_force(atom, DT)
        o.wall_post_force(wi, DT, call != 0)
        a = o.atoms()
        assert np.array_equal(atom.f.to_numpy()[:n], a["f"]), f"force, call {call}"
        assert np.array_equal(atom.torque.to_numpy()[:n], a["torque"]), f"torque, call {call}"
        sh, tc = o.wall_history(wi)
        assert np.array_equal(fix.shear.to_numpy()[:n], sh)
        assert np.array_equal(fix.touch.to_numpy()[:n], tc)
    # touch flags are kept only with history: count the atoms the wall pushes
    assert np.count_nonzero(np.any(a["f"] != 0.0, axis=1)) > 5, "the case should hold contacts"


@pytest.mark.gpu
@pytest.mark.parametrize("history", [True, False])
def test_spheres_falling_on_a_floor_bitexact_run(history):
    """hooke/history spheres thrown onto a floor wall (z fixed, x/y periodic)
    through two Simulation.run calls (20 and 180 steps; the bottom layer hits
    the floor and bounces), against the oracle's runs with the same wall:
    positions, velocities, spins, forces and the wall history."""
    import taichimps_amd as tm
    from taichimps_amd import workloads
    w = workloads.lattice(6, 6, 4, 1.1, 13, "uniform")
    n = len(w["x"])
    w["v"][:, 2] -= 2.0
    lo, hi = w["boxlo"], w["boxhi"]
    # the lowest layer (centres 0.55 above lo) starts just clear of or
    # touching the floor and is thrown onto it
    floor = float(lo[2]) + 0.08
    dom = tm.Domain(lo, hi, ("p", "p", "f"))
    atom = tm.AtomSystem(max_atoms=n)
    atom.add_particles(x=w["x"], radius=w["radius"], density=w["density"], v=w["v"],
                       omega=w["omega"], tag=w["tag"])
    nl = tm.NeighborList(dom, skin=0.1, max_atoms=n, max_neighbors_per_atom=64)
    hist = tm.ContactHistory(n, 64)
    c = dict(kn=2.0e5, gamman=50.0, kt=2.0 / 7.0 * 2.0e5, gammat=25.0, xmu=0.5)
    pair = tm.GranHookeHistory(dom, c["kn"], c["gamman"], c["kt"], c["gammat"], c["xmu"])
    sim = tm.Simulation(dom, atom, nl, hist, pair, dt=DT)
    sim.add_fix(tm.FixNVESphere(dom))
    wall = tm.FixWallGran(dom, 2, -1, floor, c["kn"], c["gamman"], c["kt"], c["gammat"],
                          c["xmu"], history=history, max_atoms=n)
    sim.add_fix(wall)

    o = O.from_workload(w)
    o.set_box(lo, hi, (1, 1, 0))
    o.set_neigh(0.1)
    o.set_pair(O.legacy_pair("hooke/history", **c))
    wi = o.add_wall(2, -1, floor, c["kn"], c["gamman"], c["kt"], c["gammat"], c["xmu"],
                    history=history)
    for steps in (20, 180):
        sim.run(steps)
        assert o.run(steps, DT) == 0
        a = o.atoms()
        for key, field in (("x", atom.x), ("v", atom.v), ("omega", atom.omega),
                           ("f", atom.f), ("torque", atom.torque)):
            assert np.array_equal(field.to_numpy()[:n], a[key]), f"{key} after {steps}"
        sh, tc = o.wall_history(wi)
        if history:
            assert np.array_equal(wall.shear.to_numpy()[:n], sh)
            assert np.array_equal(wall.touch.to_numpy()[:n], tc)
        if history and steps == 20:
            assert tc.sum() > 10, "the bottom layer should be on the floor"
        # the wall is applied inside the device-resident step
        assert sim.last_run_device_resident


# ------------------------------------------------- cylinder region walls (CPU)

def _cyl_one(x, v=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0), r=0.5):
    o = O.OracleSim(1)
    rr = np.array([r])
    o.set_atoms(np.array([x], dtype=float), np.array([v], dtype=float),
                np.array([omega], dtype=float), rr, O.mass_from_radius(rr, 1.0))
    o.set_box((-3.0,) * 3, (3.0,) * 3, (0, 0, 0))
    o.set_neigh(0.1)
    return o, O.mass_from_radius(rr, 1.0)[0]


@pytest.mark.parametrize("side", ["in", "out"])
@pytest.mark.parametrize("fstyle", ["hooke", "hertz"])
def test_cylinder_normal_force_closed_form(side, fstyle):
    """A resting sphere 0.05 into the curved side of a z-cylinder (R = 2):
    kn delta (times sqrt(r delta) for hertz) along the inward normal."""
    rp = 1.55 if side == "in" else 2.45
    c = np.array([0.3, -0.2])
    x = (c[0] + rp * 0.6, c[1] + rp * 0.8, 0.7)
    o, m = _cyl_one(x)
    w = o.add_wall_cylinder(2, c[0], c[1], 2.0, side, KN, GN, KT, GT, MU, fstyle=fstyle)
    f, t = _force(o, w)
    delta = 0.05
    fn = KN * delta * (math.sqrt(0.5 * delta) if fstyle == "hertz" else 1.0)
    sgn = -1.0 if side == "in" else 1.0
    np.testing.assert_allclose(f, [sgn * 0.6 * fn, sgn * 0.8 * fn, 0.0], rtol=1e-9, atol=1e-9)
    assert np.array_equal(t, np.zeros(3))


def test_cylinder_tangential_stick_along_the_axis_and_bounds():
    o, m = _cyl_one((1.55, 0.0, 0.0), v=(0.0, 0.0, 0.2))
    w = o.add_wall_cylinder(2, 0.0, 0.0, 2.0, "in", KN, GN, KT, GT, MU, fstyle="hertz/history")
    f, t = _force(o, w)
    poly = math.sqrt(0.5 * 0.05)
    fs = -KT * poly * 0.2 * DT - m * GT * poly * 0.2
    assert f[2] == pytest.approx(fs, rel=1e-12)
    # n = -x: torque = -r n x fs = -0.5 * (-x) x (fs z) = -0.5 * fs * y
    np.testing.assert_allclose(t, [0.0, -0.5 * fs, 0.0], rtol=1e-12)
    sh, tc = o.wall_history(w)
    assert tc[0] == 1 and sh[0][2] == pytest.approx(0.2 * DT, rel=1e-12)
    # the same sphere outside [axis_lo, axis_hi] feels nothing and is reset
    o2, _ = _cyl_one((1.55, 0.0, 0.0), v=(0.0, 0.0, 0.2))
    w2 = o2.add_wall_cylinder(2, 0.0, 0.0, 2.0, "in", KN, GN, KT, GT, MU,
                              fstyle="hooke/history", axis_lo=0.5, axis_hi=2.0)
    f2, _ = _force(o2, w2)
    assert np.array_equal(f2, np.zeros(3)) and o2.wall_history(w2)[1][0] == 0


def test_cylinder_axis_atom_and_separation_feel_nothing():
    o, _ = _cyl_one((0.0, 0.0, 0.0))
    w = o.add_wall_cylinder(2, 0.0, 0.0, 0.4, "in", KN, GN, KT, GT, MU)
    assert np.array_equal(_force(o, w)[0], np.zeros(3))
    o, _ = _cyl_one((1.0, 0.0, 0.0))
    w = o.add_wall_cylinder(2, 0.0, 0.0, 2.0, "in", KN, GN, KT, GT, MU)
    assert np.array_equal(_force(o, w)[0], np.zeros(3))


# ------------------------------------------------- cylinder region walls (GPU)

def _cyl_cases():
    out = []
    for axis in range(3):
        for side in ("in", "out"):
            out.append(dict(axis=axis, side=side, fstyle="hooke/history"))
    out += [dict(axis=2, side="in", fstyle="hertz/history"),
            dict(axis=2, side="in", fstyle="hertz"),
            dict(axis=1, side="out", fstyle="hooke", limit_damping=True),
            dict(axis=0, side="in", fstyle="hertz/history", axis_lo=1.0, axis_hi=3.0)]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case", _cyl_cases())
def test_cylinder_wall_bitexact_per_call(case):
    import taichimps_amd as tm
    x, r, v, om = _box_system(17)
    n = len(x)
    R = 1.5 if case["side"] == "in" else 1.2
    dom = tm.Domain((0.0, 0.0, 0.0), (4.0, 4.0, 4.0), ("f", "f", "f"))
    atom = tm.AtomSystem(max_atoms=n)
    atom.add_particles(x=x, radius=r, density=1.0, v=v, omega=om)
    atom._bind_domain(dom)
    kw = {k: case[k] for k in ("limit_damping", "axis_lo", "axis_hi") if k in case}
    fix = tm.FixWallGranRegion(dom, axis=case["axis"], c1=2.0, c2=2.1, radius=R,
                               side=case["side"], fstyle=case["fstyle"], kn=KN, gamman=GN,
                               kt=KT, gammat=GT, xmu=MU, max_atoms=n, **kw)
    o = O.OracleSim(n)
    o.set_atoms(x, v, om, r, O.mass_from_radius(r, 1.0))
    o.set_box((0.0,) * 3, (4.0,) * 3, (0, 0, 0))
    wi = o.add_wall_cylinder(case["axis"], 2.0, 2.1, R, case["side"], KN, GN, KT, GT, MU,
                             fstyle=case["fstyle"], **kw)
    for call in range(4):
        atom.clear_forces()
        o.clear_forces()
        fix.history_update = call != 0
        fix.post_force(atom, DT)
        o.wall_post_force(wi, DT, call != 0)
        a = o.atoms()
        assert np.array_equal(atom.f.to_numpy()[:n], a["f"]), f"force, call {call}"
        assert np.array_equal(atom.torque.to_numpy()[:n], a["torque"]), f"torque, call {call}"
        sh, tc = o.wall_history(wi)
        assert np.array_equal(fix.shear.to_numpy()[:n], sh)
        assert np.array_equal(fix.touch.to_numpy()[:n], tc)
    assert np.count_nonzero(np.any(a["f"] != 0.0, axis=1)) > 5, "the case should hold contacts"


@pytest.mark.gpu
def test_column_in_a_cylinder_on_a_floor_bitexact_run():
    """A column of spheres inside a z-cylinder container standing on a floor
    wall, thrown down and sideways: Simulation.run with both walls against
    the oracle's run with the same walls, bit for bit."""
    import taichimps_amd as tm
    from taichimps_amd import workloads
    w = workloads.lattice(4, 4, 5, 1.1, 29, "uniform")
    n = len(w["x"])
    w["v"][:, 2] -= 1.5
    w["v"][:, 0] += 0.8
    lo, hi = w["boxlo"], w["boxhi"]
    cx, cy = 0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1])
    R = 2.6
    floor = float(lo[2]) + 0.08
    c = dict(kn=2.0e5, gamman=50.0, kt=2.0e5 * 2.0 / 7.0, gammat=25.0, xmu=0.5)
    dom = tm.Domain(lo, hi, ("f", "f", "f"))
    atom = tm.AtomSystem(max_atoms=n)
    atom.add_particles(x=w["x"], radius=w["radius"], density=w["density"], v=w["v"],
                       omega=w["omega"], tag=w["tag"])
    nl = tm.NeighborList(dom, skin=0.1, max_atoms=n, max_neighbors_per_atom=64)
    sim = tm.Simulation(dom, atom, nl, tm.ContactHistory(n, 64),
                        tm.GranHookeHistory(dom, c["kn"], c["gamman"], c["kt"], c["gammat"],
                                            c["xmu"]), dt=DT)
    sim.add_fix(tm.FixNVESphere(dom))
    sim.add_fix(tm.FixWallGran(dom, 2, -1, floor, c["kn"], c["gamman"], c["kt"], c["gammat"],
                               c["xmu"], max_atoms=n))
    cyl = tm.FixWallGranRegion(dom, axis="z", c1=cx, c2=cy, radius=R, side="in",
                               fstyle="hooke/history", kn=c["kn"], gamman=c["gamman"],
                               kt=c["kt"], gammat=c["gammat"], xmu=c["xmu"], max_atoms=n)
    sim.add_fix(cyl)
    o = O.from_workload(w)
    o.set_box(lo, hi, (0, 0, 0))
    o.set_neigh(0.1)
    o.set_pair(O.legacy_pair("hooke/history", **c))
    o.add_wall(2, -1, floor, c["kn"], c["gamman"], c["kt"], c["gammat"], c["xmu"])
    wc = o.add_wall_cylinder(2, cx, cy, R, "in", c["kn"], c["gamman"], c["kt"], c["gammat"],
                             c["xmu"], fstyle="hooke/history")
    for steps in (30, 170):
        sim.run(steps)
        assert o.run(steps, DT) == 0
        a = o.atoms()
        for key, field in (("x", atom.x), ("v", atom.v), ("omega", atom.omega),
                           ("f", atom.f), ("torque", atom.torque)):
            assert np.array_equal(field.to_numpy()[:n], a[key]), f"{key} after {steps}"
        sh, tc = o.wall_history(wc)
        assert np.array_equal(cyl.shear.to_numpy()[:n], sh)
        assert np.array_equal(cyl.touch.to_numpy()[:n], tc)
        assert sim.last_run_device_resident


def _floor_sim(float_type="f64", order=("cundall", "wall"), seed=31, history=True):
    """Spheres thrown onto a floor wall (x/y periodic), nve/sphere, optional
    damping/cundall before or after the wall among the fixes."""
    import taichimps_amd as tm
    from taichimps_amd import workloads
    w = workloads.lattice(6, 5, 4, 1.1, seed, "uniform")
    n = len(w["x"])
    w["v"][:, 2] -= 2.0
    w["v"][:, 0] += 0.5
    lo, hi = w["boxlo"], w["boxhi"]
    floor = float(lo[2]) + 0.08
    c = dict(kn=2.0e5, gamman=50.0, kt=2.0 / 7.0 * 2.0e5, gammat=25.0, xmu=0.5)
    dom = tm.Domain(lo, hi, ("p", "p", "f"))
    atom = tm.AtomSystem(max_atoms=n, float_type=float_type)
    atom.add_particles(x=w["x"], radius=w["radius"], density=w["density"], v=w["v"],
                       omega=w["omega"], tag=w["tag"])
    nl = tm.NeighborList(dom, skin=0.1, max_atoms=n, max_neighbors_per_atom=64)
    pair = tm.GranHookeHistory(dom, c["kn"], c["gamman"], c["kt"], c["gammat"], c["xmu"])
    sim = tm.Simulation(dom, atom, nl, tm.ContactHistory(n, 64), pair, dt=DT)
    sim.add_fix(tm.FixNVESphere(dom))
    wall = tm.FixWallGran(dom, 2, -1, floor, c["kn"], c["gamman"], c["kt"], c["gammat"],
                          c["xmu"], history=history, shear_axis=0, shear_vel=0.3, max_atoms=n)
    for name in order:
        sim.add_fix(wall if name == "wall" else tm.FixDampingCundall(dom, 0.1, 0.1))
    return sim, wall, w, c, floor


@pytest.mark.gpu
def test_sheared_floor_with_cundall_device_run_bitexact():
    """A sheared floor wall after damping/cundall among the fixes (the
    oracle's post_force order), FP64, through two device-resident runs with
    list rebuilds: positions, velocities, spins, forces, torques and the wall
    history bit for bit with the oracle."""
    sim, wall, w, c, floor = _floor_sim()
    n = len(w["x"])
    o = O.from_workload(w)
    o.set_box(w["boxlo"], w["boxhi"], (1, 1, 0))
    o.set_neigh(0.1)
    o.set_pair(O.legacy_pair("hooke/history", **c))
    o.set_cundall(0.1, 0.1)
    wi = o.add_wall(2, -1, floor, c["kn"], c["gamman"], c["kt"], c["gammat"], c["xmu"],
                    shear_axis=0, shear_vel=0.3)
    touched = 0
    for steps in (6, 34, 160):
        sim.run(steps)
        assert sim.last_run_device_resident
        assert o.run(steps, DT) == 0
        a = o.atoms()
        for key, field in (("x", sim.atom.x), ("v", sim.atom.v), ("omega", sim.atom.omega),
                           ("f", sim.atom.f), ("torque", sim.atom.torque)):
            assert np.array_equal(field.to_numpy()[:n], a[key]), f"{key} after {steps}"
        sh, tc = o.wall_history(wi)
        assert np.array_equal(wall.shear.to_numpy()[:n], sh)
        assert np.array_equal(wall.touch.to_numpy()[:n], tc)
        touched = max(touched, int(tc.sum()))
    assert touched > 5, "the bottom layer should reach the floor"
    assert sim.neighbor.build_count >= 2


@pytest.mark.gpu
@pytest.mark.parametrize("order", [("wall", "cundall"), ("cundall", "wall")])
def test_wall_order_in_device_run_equals_stage_steps(order):
    """The wall before or after damping/cundall: the device-resident run
    equals the reference's stage sequence (Simulation.run with pre yes:
    init_simulation, then step() per step, post_force in the fixes' order)
    bit for bit in FP64, over two runs."""
    sim, wall, w, _, _ = _floor_sim(order=order)
    ref, rwall, _, _, _ = _floor_sim(order=order)
    n = len(w["x"])
    touched = 0
    for k, steps in enumerate((6, 114)):
        sim.run(steps)
        assert sim.last_run_device_resident
        ref.setup_run(steps)
        ref.init_simulation()
        for _ in range(steps):
            ref.step()
        for key in ("x", "v", "omega", "f", "torque"):
            got = getattr(sim.atom, key).to_numpy()[:n]
            want = getattr(ref.atom, key).to_numpy()[:n]
            assert np.array_equal(got, want), f"{key} after {steps}"
        assert np.array_equal(wall.shear.to_numpy()[:n], rwall.shear.to_numpy()[:n])
        touched = max(touched, int(rwall.touch.to_numpy()[:n].sum()))
    assert touched > 5, "the bottom layer should reach the floor"


@pytest.mark.gpu
def test_fp32_split_floor_device_run_tracks_fp64_oracle():
    """FP32-split, the floor wall inside the device-resident step: positions
    within 1e-4 and velocities within 1e-3 of the largest speed of the FP64
    oracle after 150 steps; the same atoms rest on the floor."""
    sim, wall, w, c, floor = _floor_sim(float_type="f32split")
    n = len(w["x"])
    o = O.from_workload(w)
    o.set_box(w["boxlo"], w["boxhi"], (1, 1, 0))
    o.set_neigh(0.1)
    o.set_pair(O.legacy_pair("hooke/history", **c))
    o.set_cundall(0.1, 0.1)
    wi = o.add_wall(2, -1, floor, c["kn"], c["gamman"], c["kt"], c["gammat"], c["xmu"],
                    shear_axis=0, shear_vel=0.3)
    sim.run(150)
    assert sim.last_run_device_resident
    assert o.run(150, DT) == 0
    a = o.atoms()
    x = sim.atom.x.to_numpy()[:n]
    v = sim.atom.v.to_numpy()[:n]
    assert np.abs(x - a["x"]).max() < 1e-4
    assert np.abs(v - a["v"]).max() < 1e-3 * np.abs(a["v"]).max()
    _, tc = o.wall_history(wi)
    assert np.array_equal(wall.touch.to_numpy()[:n], tc)


@pytest.mark.gpu
@pytest.mark.parametrize("geometry", ["plane", "cylinder"])
def test_walls_on_fp32_split_systems_within_tolerance(geometry):
    """FP32-split systems: velocities, spins, forces and torques in FP32, the
    contact in FP64 from decoded positions. One call against the FP64
    oracle: forces within 1e-4 of the largest wall force (positions carry the
    split encoding's FP32 rounding); the same atoms touch."""
    import taichimps_amd as tm
    x, r, v, om = _box_system(23)
    n = len(x)
    dom = tm.Domain((0.0, 0.0, 0.0), (4.0, 4.0, 4.0), ("f", "f", "f"))
    atom = tm.AtomSystem(max_atoms=n, float_type="f32split")
    atom.add_particles(x=x, radius=r, density=1.0, v=v, omega=om)
    atom._bind_domain(dom)
    o = O.OracleSim(n)
    o.set_atoms(x, v, om, r, O.mass_from_radius(r, 1.0))
    o.set_box((0.0,) * 3, (4.0,) * 3, (0, 0, 0))
    if geometry == "plane":
        fix = tm.FixWallGran(dom, 2, -1, 0.25, KN, GN, KT, GT, MU, max_atoms=n)
        wi = o.add_wall(2, -1, 0.25, KN, GN, KT, GT, MU)
    else:
        fix = tm.FixWallGranRegion(dom, axis="z", c1=2.0, c2=2.1, radius=1.5, side="in",
                                   fstyle="hertz/history", kn=KN, gamman=GN, kt=KT, gammat=GT,
                                   xmu=MU, max_atoms=n)
        wi = o.add_wall_cylinder(2, 2.0, 2.1, 1.5, "in", KN, GN, KT, GT, MU,
                                 fstyle="hertz/history")
    atom.clear_forces()
    o.clear_forces()
    fix.post_force(atom, DT)
    o.wall_post_force(wi, DT, True)
    a = o.atoms()
    scale = np.abs(a["f"]).max()
    assert scale > 0.0
    assert np.abs(atom.f.to_numpy()[:n] - a["f"]).max() < 1e-4 * scale
    assert np.abs(atom.torque.to_numpy()[:n] - a["torque"]).max() < 1e-4 * max(
        np.abs(a["torque"]).max(), 1e-30)
    assert np.array_equal(fix.touch.to_numpy()[:n], o.wall_history(wi)[1])
