"""CPU pins of the oracle's PairGranular contact law (reference
src/taichimps/GRANULAR/granular.py:282-430, coefficients :94-269).

Every check here is a closed form of contact mechanics, a textbook
relation or a physical consequence (the restitution a damping model is
built to give), not the oracle's own expression order. Two spheres sit on
the x axis: atom 0 at the origin side, atom 1 at distance L = r0 + r1 - d,
so the unit normal from 1 to 0 is n = (-1, 0, 0) and the force reported on
atom 0 is fntot * n + fs.

A plausible restatement slip fails at least one test: a dropped contact
radius in viscoelastic damping, r instead of r - d/2 as the moment arm, a
missing sqrt in tsuji/coeff_restitution damping, an unscaled mindlin
stiffness, a history kept with its normal part, a slip history that keeps
the damping share, the viscous cap of linear_nohistory.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O

R0, R1, D = 0.5, 0.4, 0.02
E, NU = 1.0e7, 0.3
DT = 1.0e-4


def _two(pair, v0=(0.0, 0.0, 0.0), v1=(0.0, 0.0, 0.0), r0=R0, r1=R1, d=D, w0=None, w1=None):
    x = np.array([[3.0, 3.0, 3.0], [3.0 + r0 + r1 - d, 3.0, 3.0]])
    r = np.array([r0, r1])
    o = O.OracleSim(2, max_nb=8)
    om = np.zeros((2, 3))
    if w0 is not None:
        om[0] = w0
    if w1 is not None:
        om[1] = w1
    o.set_atoms(x, np.array([v0, v1], dtype=np.float64), om, r, O.mass_from_radius(r, 1.0))
    o.set_box((0.0,) * 3, (10.0,) * 3, (1, 1, 1))
    o.set_neigh(0.1)
    o.set_pair(pair)
    assert o.nlist_build() == 0
    return o


def _call(o, dt=DT, shearupdate=True):
    o.clear_forces()
    o.pair_compute(dt, shearupdate)
    return o.atoms()


def _geom(r0=R0, r1=R1, d=D):
    """Effective radius, effective mass (density 1) and Hertz contact radius."""
    m0 = 4.0 / 3.0 * math.pi * r0**3
    m1 = 4.0 / 3.0 * math.pi * r1**3
    reff = r0 * r1 / (r0 + r1)
    meff = m0 * m1 / (m0 + m1)
    return reff, meff, math.sqrt(d * reff)


def _estar(e=E, nu=NU):
    """Textbook effective modulus of two identical spheres: 1/E* = 2 (1 - nu^2) / E."""
    return e / (2.0 * (1.0 - nu * nu))


def _gstar(e=E, nu=NU):
    """Textbook effective shear modulus: G = E / (2 (1 + nu)),
    1/G* = 2 (2 - nu) / G for two identical spheres."""
    g = e / (2.0 * (1.0 + nu))
    return g / (2.0 * (2.0 - nu))


# ------------------------------------------------------------ normal laws

def test_static_hertz_material_pair_is_the_hertz_law():
    """F = 4/3 E* sqrt(Reff) d^(3/2) for two identical elastic spheres at rest."""
    o = _two(O.granular_pair("hertz/material", [E, 0.5, NU], damping="viscoelastic"))
    a = _call(o)
    reff, _, _ = _geom()
    f = 4.0 / 3.0 * _estar() * math.sqrt(reff) * D**1.5
    np.testing.assert_allclose(a["f"][0], [-f, 0.0, 0.0], rtol=1e-12, atol=1e-12 * f)
    np.testing.assert_allclose(a["f"][1], [f, 0.0, 0.0], rtol=1e-12, atol=1e-12 * f)
    # no tangential force, no torque
    assert np.all(a["torque"] == 0.0)


@pytest.mark.parametrize("normal,expect", [
    ("hooke", lambda k, reff: k * D),
    ("hertz", lambda k, reff: k * D * math.sqrt(D * reff)),
])
def test_static_hooke_and_hertz_normal(normal, expect):
    k = 3.0e5
    o = _two(O.granular_pair(normal, [k, 0.0], damping="none"))
    a = _call(o)
    reff, _, _ = _geom()
    np.testing.assert_allclose(a["f"][0, 0], -expect(k, reff), rtol=1e-12)


# -------------------------------------------------------------- damping

def _approaching(normal, coeffs, damping, u=0.7):
    """Atom 0 moves towards atom 1 at u: vn = v.n = -u, so damping adds to fne."""
    o = _two(O.granular_pair(normal, coeffs, damping=damping), v0=(u, 0.0, 0.0))
    a = _call(o)
    return -a["f"][0, 0], u


def test_damping_velocity_and_mass_velocity():
    k, g = 2.0e5, 40.0
    reff, meff, a = _geom()
    fn, u = _approaching("hooke", [k, g], "velocity")
    np.testing.assert_allclose(fn, k * D + g * u, rtol=1e-12)
    fn, u = _approaching("hooke", [k, g], "mass_velocity")
    np.testing.assert_allclose(fn, k * D + g * meff * u, rtol=1e-12)


def test_damping_viscoelastic_scales_with_contact_radius():
    """viscoelastic: Fd = damp meff a vn with a = sqrt(d Reff) (the contact radius)."""
    g = 40.0
    reff, meff, a = _geom()
    fn, u = _approaching("hertz/material", [E, g, NU], "viscoelastic")
    fne = 4.0 / 3.0 * _estar() * math.sqrt(reff) * D**1.5
    np.testing.assert_allclose(fn, fne + g * meff * a * u, rtol=1e-12)


def _zeta(e):
    """Damping ratio of the linear spring-dashpot whose restitution is e."""
    le = math.log(e)
    return -le / math.sqrt(math.pi**2 + le * le)


def test_damping_coeff_restitution_hooke_closed_form():
    """Linear spring-dashpot: c = 2 zeta sqrt(meff k)."""
    k, e = 2.0e5, 0.6
    _, meff, _ = _geom()
    fn, u = _approaching("hooke", [k, e], "coeff_restitution")
    c = 2.0 * _zeta(e) * math.sqrt(meff * k)
    np.testing.assert_allclose(fn, k * D + c * u, rtol=1e-12)


def test_damping_coeff_restitution_hertz_is_tsujis_form():
    """Hertz: eta = 2 sqrt(5/6) beta sqrt(S_n meff), S_n = 2 E* a (Tsuji 1992;
    beta = -ln e / sqrt(ln^2 e + pi^2))."""
    e = 0.6
    reff, meff, a = _geom()
    fn, u = _approaching("hertz/material", [E, e, NU], "coeff_restitution")
    es = _estar()
    eta = 2.0 * math.sqrt(5.0 / 6.0) * _zeta(e) * math.sqrt(2.0 * es * a * meff)
    fne = 4.0 / 3.0 * es * math.sqrt(reff) * D**1.5
    np.testing.assert_allclose(fn, fne + eta * u, rtol=1e-12)


def test_damping_tsuji_prefactor_is_sqrt_of_meff_times_stiffness():
    """tsuji: Fd = alpha(e) sqrt(2) sqrt(meff fne / d) vn; the stiffness
    fne / d is k for hooke and k a for hertz, so the ratio of the two damping
    forces at the same e is sqrt(a) (alpha(e) cancels)."""
    k, e, u = 2.0e5, 0.7, 0.7
    reff, meff, a = _geom()
    fn_hooke, _ = _approaching("hooke", [k, e], "tsuji", u)
    fn_hertz, _ = _approaching("hertz", [k, e], "tsuji", u)
    fd_hooke = fn_hooke - k * D
    fd_hertz = fn_hertz - k * a * D
    np.testing.assert_allclose(fd_hertz / fd_hooke, math.sqrt(a), rtol=1e-12)
    # and the hooke prefactor is linear in vn and in sqrt(meff k)
    fn2, _ = _approaching("hooke", [4.0 * k, e], "tsuji", 2.0 * u)
    np.testing.assert_allclose(fn2 - 4.0 * k * D, 4.0 * fd_hooke, rtol=1e-12)


def _collide(pair, u=0.5, dt=2.0e-6, steps=6000, r=0.5):
    """Head-on collision of two equal spheres in a closed box through the
    oracle's whole step (nve/sphere + neighbor rebuilds); returns the
    restitution coefficient -(v1 - v0)' / (v1 - v0) (approach: v1 - v0 = -2u)."""
    gap = 1.0e-6
    x = np.array([[4.0, 5.0, 5.0], [4.0 + 2 * r + gap, 5.0, 5.0]])
    v = np.array([[u, 0.0, 0.0], [-u, 0.0, 0.0]])
    rr = np.array([r, r])
    o = O.OracleSim(2, max_nb=8)
    o.set_atoms(x, v, np.zeros((2, 3)), rr, O.mass_from_radius(rr, 1.0))
    o.set_box((0.0,) * 3, (10.0,) * 3, (0, 0, 0))
    o.set_neigh(0.05)
    o.set_pair(pair)
    assert o.run(steps, dt) == 0
    a = o.atoms()
    sep = a["x"][1, 0] - a["x"][0, 0]
    assert sep > 2 * r, "the spheres are still in contact"
    return (a["v"][1, 0] - a["v"][0, 0]) / (2.0 * u)


@pytest.mark.parametrize("e", [0.3, 0.6, 0.9])
def test_coeff_restitution_hooke_gives_that_restitution(e):
    """A linear spring-dashpot with c = 2 zeta sqrt(m k) separates (overlap
    back to zero) with relative speed e times the approach speed, exactly in
    continuous time: the damped half-period ends at e = exp(-zeta pi / sqrt(1 - zeta^2))."""
    cor = _collide(O.granular_pair("hooke", [2.0e5, e], damping="coeff_restitution"))
    assert abs(cor - e) < 2e-3, cor


@pytest.mark.parametrize("e", [0.5, 0.8])
def test_tsuji_damping_gives_about_that_restitution(e):
    """Tsuji's Hertzian damping is fitted (Marshall 2009, eq. 53) so that the
    restitution of a Hertzian impact is e, independent of speed."""
    pair = O.granular_pair("hertz/material", [E, e, NU], damping="tsuji")
    c1 = _collide(pair, u=0.5, steps=8000)
    c2 = _collide(pair, u=0.2, steps=12000)
    assert abs(c1 - e) < 0.02 * e, c1
    assert abs(c1 - c2) < 2e-3, (c1, c2)


def test_limit_damping_clamps_an_attractive_normal_force():
    """Separating fast with strong damping: fntot = fne - pref vn < 0 becomes 0."""
    k, g, u = 2.0e5, 2.0e5, 1.0
    free = _two(O.granular_pair("hooke", [k, g], damping="velocity"), v0=(-u, 0.0, 0.0))
    clamp = _two(O.granular_pair("hooke", [k, g], damping="velocity", limit_damping=True),
                 v0=(-u, 0.0, 0.0))
    a = _call(free)
    np.testing.assert_allclose(-a["f"][0, 0], k * D - g * u, rtol=1e-12)
    b = _call(clamp)
    assert np.all(b["f"] == 0.0)


# ------------------------------------------------------------ tangential

def _sliding(tangential, tcoeffs, u, damping="viscoelastic", g=40.0, calls=1, mu=None):
    """Atom 1 slides along +y at u: vtr = v0 - v1 = (0, -u, 0), vn = 0."""
    pair = O.granular_pair("hertz/material", [E, g, NU], tangential=tangential,
                           tangential_coeffs=tcoeffs, damping=damping)
    o = _two(pair, v1=(0.0, u, 0.0))
    for _ in range(calls):
        a = _call(o)
    return a, o


def test_mindlin_stick_accumulates_history_and_scales_stiffness_by_contact_radius():
    """k calls below slip: h = sum vtr dt, Fs = -kt a h - xt (damp meff a) vtr,
    with kt = 8 G* (NULL); torque on each sphere (r - d/2) Fs about z."""
    g, xt, mu, u, k = 40.0, 0.5, 0.5, 0.01, 3
    a, o = _sliding("mindlin", [None, xt, mu], u, g=g, calls=k)
    reff, meff, ac = _geom()
    kt = 8.0 * _gstar()
    fs = kt * ac * (k * DT * u) + xt * (g * meff * ac) * u
    fne = 4.0 / 3.0 * _estar() * math.sqrt(reff) * D**1.5
    assert fs < mu * fne  # stick
    np.testing.assert_allclose(a["f"][0, 1], fs, rtol=1e-12)
    np.testing.assert_allclose(a["f"][1, 1], -fs, rtol=1e-12)
    np.testing.assert_allclose(a["torque"][0, 2], (R0 - 0.5 * D) * fs, rtol=1e-12)
    np.testing.assert_allclose(a["torque"][1, 2], (R1 - 0.5 * D) * fs, rtol=1e-12)
    _, _, _, shear = o.pairs()
    np.testing.assert_allclose(shear[0], [0.0, -k * DT * u, 0.0], rtol=1e-12, atol=1e-300)


def test_linear_history_stiffness_is_not_scaled():
    g, xt, mu, u, k, kt = 40.0, 0.5, 0.5, 0.01, 2, 1.0e5
    a, _ = _sliding("linear_history", [kt, xt, mu], u, g=g, calls=k)
    _, meff, ac = _geom()
    fs = kt * (k * DT * u) + xt * (g * meff * ac) * u
    np.testing.assert_allclose(a["f"][0, 1], fs, rtol=1e-12)


def test_equal_spheres_moment_arm_is_half_the_centre_distance():
    """For equal radii the contact point is midway: arm = L / 2 = r - d/2."""
    pair = O.granular_pair("hertz/material", [E, 40.0, NU], tangential="mindlin",
                           tangential_coeffs=[None, 0.5, 0.5], damping="viscoelastic")
    o = _two(pair, v1=(0.0, 0.01, 0.0), r0=0.5, r1=0.5, d=0.04)
    a = _call(o)
    half = (0.5 + 0.5 - 0.04) / 2.0
    np.testing.assert_allclose(a["torque"][0, 2], half * a["f"][0, 1], rtol=1e-12)


def test_mindlin_slip_caps_at_coulomb_and_keeps_the_elastic_share():
    """At slip |Ft| = mu |Fn| and the stored history is the elastic share of
    the capped force: h = -(Fs - Fdamp) / (kt a)."""
    g, xt, mu, u = 40.0, 0.5, 0.3, 100.0
    a, o = _sliding("mindlin", [None, xt, mu], u, g=g, calls=1)
    reff, meff, ac = _geom()
    fne = 4.0 / 3.0 * _estar() * math.sqrt(reff) * D**1.5
    f = a["f"][0]
    n = np.array([-1.0, 0.0, 0.0])
    fn = f @ n
    ft = f - fn * n
    np.testing.assert_allclose(fn, fne, rtol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(ft), mu * fne, rtol=1e-12)
    fdamp = xt * (g * meff * ac) * u
    assert fdamp < mu * fne  # the elastic share is positive
    kt = 8.0 * _gstar()
    _, _, _, shear = o.pairs()
    np.testing.assert_allclose(shear[0, 1], -(mu * fne - fdamp) / (kt * ac), rtol=1e-12)
    np.testing.assert_allclose(shear[0, [0, 2]], [0.0, 0.0], atol=1e-300)


def test_rotate_rescale_keeps_the_magnitude_and_drops_the_normal_part():
    """A stored history with a normal component is projected onto the tangent
    plane and rescaled to its old length before the update."""
    pair = O.granular_pair("hertz/material", [E, 40.0, NU], tangential="mindlin",
                           tangential_coeffs=[None, 0.5, 0.5], damping="viscoelastic")
    o = _two(pair)
    h0 = np.array([3.0e-6, -4.0e-6, 0.0])  # normal part along x
    o.set_history([2], h0.reshape(1, 3))
    a = _call(o)
    _, _, _, shear = o.pairs()
    h = shear[0]
    np.testing.assert_allclose(np.linalg.norm(h), np.linalg.norm(h0), rtol=1e-12)
    assert abs(h[0]) < 1e-20
    np.testing.assert_allclose(h, [0.0, -5.0e-6, 0.0], rtol=1e-12, atol=1e-20)
    # the force uses the rotated history
    reff, _, ac = _geom()
    kt = 8.0 * _gstar()
    np.testing.assert_allclose(a["f"][0, 1], kt * ac * 5.0e-6, rtol=1e-12)
    # without the history update neither rotation nor increment happens
    o2 = _two(pair)
    o2.set_history([2], h0.reshape(1, 3))
    _call(o2, shearupdate=False)
    _, _, _, shear2 = o2.pairs()
    np.testing.assert_array_equal(shear2[0], h0)


def test_history_is_reset_for_a_new_partner():
    """A stored history whose partner tag is not j's is discarded."""
    pair = O.granular_pair("hertz/material", [E, 40.0, NU], tangential="mindlin",
                           tangential_coeffs=[None, 0.5, 0.5], damping="viscoelastic")
    o = _two(pair, v1=(0.0, 0.01, 0.0))
    o.set_history([7], np.array([[0.0, 1.0e-3, 0.0]]))
    _call(o)
    pi, pj, pa, shear = o.pairs()
    assert pa[0] == 2
    np.testing.assert_allclose(shear[0], [0.0, -DT * 0.01, 0.0], rtol=1e-12)


def test_rolling_spin_enters_the_tangential_velocity():
    """vtr = vt - (r0 w0 + r1 w1) x n: equal spins about z, no translation,
    with n = -x give vtr = (r0 + r1) w (0, -1, 0)... (x cross z) sign fixed by hand."""
    w = 0.3
    pair = O.granular_pair("hertz/material", [E, 40.0, NU], tangential="mindlin",
                           tangential_coeffs=[None, 0.5, 0.5], damping="viscoelastic")
    o = _two(pair, w0=(0.0, 0.0, w), w1=(0.0, 0.0, w))
    _call(o)
    _, _, _, shear = o.pairs()
    # wr = (R0 + R1) w z; wr x n = (R0 + R1) w (z x -x) = -(R0 + R1) w y; vtr = +(R0 + R1) w y
    np.testing.assert_allclose(shear[0], [0.0, DT * (R0 + R1) * w, 0.0], rtol=1e-12)


def test_linear_nohistory_is_viscous_below_the_cap_and_coulomb_above():
    g, xt, mu = 40.0, 0.5, 0.5
    reff, meff, ac = _geom()
    fne = 4.0 / 3.0 * _estar() * math.sqrt(reff) * D**1.5
    damp_t = xt * g * meff * ac
    u = 0.01
    a, _ = _sliding("linear_nohistory", [xt, mu], u, g=g)
    assert damp_t * u < mu * fne
    np.testing.assert_allclose(a["f"][0, 1], damp_t * u, rtol=1e-12)
    u = 1.0e5
    a, _ = _sliding("linear_nohistory", [xt, mu], u, g=g)
    assert damp_t * u > mu * fne
    np.testing.assert_allclose(a["f"][0, 1], mu * fne, rtol=1e-12)


# ------------------------------- reference tests/test_additional_models.py

def test_pair_granular_and_wall_from_the_reference_suite():
    """tests/test_additional_models.py:70-97: one sphere (r 0.1, density
    1000) at x = 0.05 moving at -1 towards a lower x wall at 0 is pushed
    back (kn d = 1e4 * 0.05 with gamman 0); PairGranular hooke with
    mass_velocity damping on that lone sphere adds nothing."""
    r = np.array([0.1])
    o = O.OracleSim(1, max_nb=5)
    o.set_atoms(np.array([[0.05, 0.5, 0.5]]), np.array([[-1.0, 0.0, 0.0]]), np.zeros((1, 3)), r,
                O.mass_from_radius(r, 1000.0))
    o.set_box((0.0,) * 3, (10.0,) * 3, (1, 1, 1))
    o.set_neigh(0.1)
    w = o.add_wall(0, -1, 0.0, 1e4, 0.0, 0.0, 0.0, 0.0)
    o.clear_forces()
    o.wall_post_force(w, 0.001)
    f = o.atoms()["f"]
    assert f[0, 0] > 0.0
    np.testing.assert_allclose(f[0, 0], 1e4 * 0.05, rtol=1e-12)
    o.set_pair(O.granular_pair("hooke", [1000.0, 0.0], damping="mass_velocity"))
    assert o.nlist_build() == 0
    o.clear_forces()
    o.pair_compute(0.001)
    assert np.all(o.atoms()["f"] == 0.0)
