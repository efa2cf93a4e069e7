"""CPU pins of the oracle's fix deform/pressure style table (reference
src/taichimps/EXTRA_FIX/deform_pressure.py:582-702 device form,
:758-857 host form) and of the `remap v` partner-velocity shift
(src/taichimps/domain.py:388-414).

Each box length is checked against the closed form the style defines
(L0 + v t, L0 + A sin(2 pi t / P), linear interpolation to the final,
delta or scaled box over the run), each pressure update against one
update worked by hand from a kinetic-only pressure tensor, and the
velocity shift against the motion of a periodic image of a deforming box.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O

L = 10.0


def _box_sim(v=None, x=None, n=4, periodic=(1, 1, 1), radius=0.1):
    rng = np.random.default_rng(11)
    x = rng.uniform(1.0, 9.0, size=(n, 3)) if x is None else np.asarray(x, dtype=np.float64)
    n = len(x)
    v = np.zeros((n, 3)) if v is None else np.asarray(v, dtype=np.float64)
    r = np.full(n, radius)
    o = O.OracleSim(n, max_nb=8)
    o.set_atoms(x, v, np.zeros((n, 3)), r, O.mass_from_radius(r, 1.0))
    o.set_box((0.0,) * 3, (L,) * 3, periodic)
    o.set_neigh(0.1)
    return o, x, v, r


def _steps(o, k, dt):
    for _ in range(k):
        assert o.servo_end_of_step(dt) == 0
    return o.box()


@pytest.mark.parametrize("k", [1, 7, 40])
def test_vel_and_wiggle_box_lengths(k):
    """vel: L(t) = L0 + vel t; wiggle: L(t) = L0 + A sin(2 pi t / P), both
    symmetric about the start centre; t = k dt after k updates."""
    dt, vel, amp, per = 1e-3, -0.8, 0.3, 0.05
    o, _, _, _ = _box_sim()
    o.set_servo(O.servo_params(axes={"x": {"style": "vel", "vel": vel},
                                     "y": {"style": "wiggle", "amplitude": amp, "tperiod": per},
                                     "z": {"style": "none"}}))
    o.servo_setup(100, dt)
    lo, hi, _ = _steps(o, k, dt)
    t = k * dt
    np.testing.assert_allclose(hi[0] - lo[0], L + vel * t, rtol=1e-13)
    np.testing.assert_allclose(hi[1] - lo[1], L + amp * math.sin(2.0 * math.pi * t / per),
                               rtol=1e-13)
    np.testing.assert_allclose(0.5 * (lo[:2] + hi[:2]), [0.5 * L] * 2, rtol=1e-13)
    np.testing.assert_array_equal([lo[2], hi[2]], [0.0, L])


@pytest.mark.parametrize("k", [3, 25])
def test_final_delta_scale_interpolate_over_the_run(k):
    """final/delta/scale move the faces linearly in k / nsteps_total to the
    final box [flo, fhi], to [lo0 + dlo, hi0 + dhi], and to the box scaled
    by s about its centre; at k = nsteps_total they reach them."""
    dt, total = 1e-3, 25
    flo, fhi, dlo, dhi, sc = -1.0, 12.0, 0.5, -1.5, 0.8
    o, _, _, _ = _box_sim()
    o.set_servo(O.servo_params(axes={"x": {"style": "final", "flo": flo, "fhi": fhi},
                                     "y": {"style": "delta", "dlo": dlo, "dhi": dhi},
                                     "z": {"style": "scale", "scale": sc}}))
    o.servo_setup(total, dt)
    lo, hi, _ = _steps(o, k, dt)
    f = k / total
    np.testing.assert_allclose([lo[0], hi[0]], [f * flo, L + f * (fhi - L)], rtol=1e-13,
                               atol=1e-13)
    np.testing.assert_allclose([lo[1], hi[1]], [f * dlo, L + f * dhi], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(hi[2] - lo[2], L * (1.0 + f * (sc - 1.0)), rtol=1e-13)
    np.testing.assert_allclose(0.5 * (lo[2] + hi[2]), 0.5 * L, rtol=1e-13)


def _kinetic_p(v, r):
    m = 4.0 / 3.0 * math.pi * r**3
    return np.array([(m * v[:, a] * v[:, a]).sum() for a in range(3)]) / L**3


@pytest.mark.parametrize("couple,groups", [
    ("xy", ((0, 1), (2,))), ("yz", ((1, 2), (0,))), ("xz", ((0, 2), (1,))),
    ("none", ((0,), (1,), (2,))), ("xyz", ((0, 1, 2),)),
])
def test_pressure_couple_averages_the_coupled_axes(couple, groups):
    """One pressure update: axes in one couple group see the group's mean
    diagonal pressure; cum_d = dt pgain (p_d - target) L, faces move by cum/2."""
    dt, pt, pg = 1e-3, 0.2, 1e-2
    v = np.array([[3.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, 1.0], [1.0, 1.0, 1.0]])
    o, x, v, r = _box_sim(v=v)
    o.set_servo(O.servo_params(axes={d: {"style": "pressure", "ptarget": pt, "pgain": pg}
                                     for d in "xyz"}, couple=couple))
    o.servo_setup(0, dt)
    lo, hi, _ = _steps(o, 1, dt)
    pdiag = _kinetic_p(v, r)
    pseen = np.zeros(3)
    for g in groups:
        pseen[list(g)] = pdiag[list(g)].mean()
    cum = dt * pg * (pseen - pt) * L
    np.testing.assert_allclose(lo, -0.5 * cum, rtol=1e-12)
    np.testing.assert_allclose(hi - L, 0.5 * cum, rtol=1e-9)


def test_max_rate_clips_and_normalize_divides_by_the_target():
    """max/rate caps |rate| (a strain rate); normalize/pressure divides the
    raw rate by |target| first; a zero target under normalize takes
    +-max_rate."""
    dt = 1e-3
    v = np.array([[3.0, 0.0, 0.0], [0.0, 2.0, 0.0], [0.0, 0.0, 1.0], [1.0, 1.0, 1.0]])
    o, x, v, r = _box_sim(v=v)
    p = _kinetic_p(v, r)
    axes = {"x": {"style": "pressure", "ptarget": 0.1, "pgain": 1.0},
            "y": {"style": "pressure", "ptarget": 0.5, "pgain": 1e-2},
            "z": {"style": "pressure", "ptarget": 0.0, "pgain": 1e-2}}
    o.set_servo(O.servo_params(axes=axes, couple="none", max_rate=0.05, normalize=True))
    o.servo_setup(0, dt)
    lo, hi, _ = _steps(o, 1, dt)
    rate = np.array([(p[0] - 0.1) / 0.1, 1e-2 * (p[1] - 0.5) / 0.5, math.copysign(0.05, p[2])])
    rate[0] = math.copysign(min(abs(rate[0]), 0.05), rate[0])
    assert abs(rate[1]) < 0.05 and abs((p[0] - 0.1) / 0.1) > 0.05
    cum = dt * rate * L
    np.testing.assert_allclose(hi - lo, L + cum, rtol=1e-12)
    # without normalize and max/rate the same update is the raw rate
    o2, _, _, _ = _box_sim(v=v)
    o2.set_servo(O.servo_params(axes=axes, couple="none"))
    o2.servo_setup(0, dt)
    lo2, hi2, _ = _steps(o2, 1, dt)
    raw = np.array([1.0 * (p[0] - 0.1), 1e-2 * (p[1] - 0.5), 1e-2 * p[2]])
    np.testing.assert_allclose(hi2 - lo2, L + dt * raw * L, rtol=1e-12)


@pytest.mark.parametrize("remap", ["x", "v", "none"])
def test_remap_modes_move_atoms_only_under_remap_x(remap):
    """remap x keeps fractional positions through a box change; v and none
    leave positions where they are (v shifts velocities of images instead)."""
    dt, rate = 1e-3, 2.0
    o, x, _, _ = _box_sim()
    o.set_servo(O.servo_params(axes={d: {"style": "erate", "rate": rate} for d in "xyz"},
                               remap=remap))
    o.servo_setup(0, dt)
    lo, hi, hr = _steps(o, 3, dt)
    xn = o.atoms()["x"]
    np.testing.assert_allclose(hi - lo, L * (1.0 + rate * 3 * dt), rtol=1e-13)
    if remap == "x":
        np.testing.assert_allclose((xn - lo) / (hi - lo), x / L, atol=1e-14)
    else:
        np.testing.assert_array_equal(xn, x)
    # h_rate is the box-length rate of the last update: rate * L0 for erate
    np.testing.assert_allclose(hr[:3], rate * L, rtol=1e-9)


def test_remap_v_shifts_the_image_partners_velocity():
    """Two resting spheres in contact across the periodic x boundary. Under
    remap v the image of atom 1 moves with the lattice: its velocity is
    raised by n * Ldot (n = image count), so the pair separates at Ldot
    and the dashpot pulls with meff gamman Ldot; remap x sees no motion."""
    r, d, ldot = 0.5, 0.02, 3.0
    kn, gn = 2.0e5, 50.0
    x = np.array([[L - r + 0.5 * d, 5.0, 5.0], [r - 0.5 * d, 5.0, 5.0]])
    pair = O.legacy_pair("hooke/history", kn=kn, gamman=gn, kt=0.0, gammat=0.0, xmu=0.0)
    out = {}
    for vremap in (0, 1):
        o, _, _, _ = _box_sim(x=x, radius=r)
        o.set_pair(pair)
        o.set_h_rate([ldot, 0.0, 0.0, 0.0, 0.0, 0.0], vremap)
        assert o.nlist_build() == 0
        o.clear_forces()
        o.pair_compute(1e-4)
        out[vremap] = o.atoms()["f"]
    m = 4.0 / 3.0 * math.pi * r**3
    # atom 0 (near hi) is pushed towards -x by the overlap; the image of atom 1
    # at x1 + L recedes at +Ldot, so the dashpot pulls atom 0 towards +x
    np.testing.assert_allclose(out[0][0], [-kn * d, 0.0, 0.0], rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(out[1][0], [-kn * d + 0.5 * m * gn * ldot, 0.0, 0.0], rtol=1e-12,
                               atol=1e-9)


def _ulps(a, b):
    """Distance in units in the last place of b."""
    return abs(a - b) / math.ulp(b) if b != 0.0 else abs(a) / 5e-324


def test_servo_sin_and_exp_are_within_an_ulp_of_libm():
    """The wiggle and trate styles evaluate sin and exp by one fixed method
    on host and device (so both round alike); it stays within two ulps of
    the C library's values, and equals them for most arguments, over the
    arguments a run meets (2 pi t / period up to thousands of periods,
    rate * t up to +-50), and is exact at the special points."""
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(-7.0, 7.0, 4000), rng.uniform(-2e4, 2e4, 4000),
                         np.arange(-64, 65) * (math.pi / 4.0)])
    u = np.array([_ulps(O.det_sin(x), math.sin(x)) for x in xs if abs(math.sin(x)) > 1e-12])
    assert u.max() <= 2.0 and np.mean(u == 0.0) > 0.75, (u.max(), np.mean(u == 0.0))
    es = np.concatenate([rng.uniform(-1.0, 1.0, 4000), rng.uniform(-50.0, 50.0, 4000)])
    u = np.array([_ulps(O.det_exp(x), math.exp(x)) for x in es])
    assert u.max() <= 1.0 and np.mean(u == 0.0) > 0.75, (u.max(), np.mean(u == 0.0))
    assert O.det_sin(0.0) == 0.0 and O.det_exp(0.0) == 1.0
    # near multiples of pi the absolute error stays at the rounding of the argument
    for k in range(1, 2000, 37):
        x = k * math.pi
        assert abs(O.det_sin(x) - math.sin(x)) < 1e-15 * k
