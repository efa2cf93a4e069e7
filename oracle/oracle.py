"""ctypes wrapper of the CPU FP64 parity oracle -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker. The product
package (``taichimps_amd``) never imports it.

The oracle restates the reference's per-step DEM path (``dem_oracle.cpp``).
This module adds the two host-side derivations the reference performs in
Python before any kernel runs, restated independently of the product code so
that the product's own derivations are checked against them:

* the pair-style coefficients (hooke_history.py:48-68, hertz_history.py:36-55,
  granular.py:94-269, input.py:444-472 for the NULL shorthands);
* per-atom mass from radius and density (atom.py:110-111).
"""

from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
# the threaded build (benchmark CPU baseline only; its sums depend on thread timing)
_OMP_LIB_PATH = os.path.join(_HERE, "liboracle_omp.so")


class PairParams(C.Structure):
    _fields_ = [
        ("style", C.c_int), ("limit_damping", C.c_int),
        ("kn", C.c_double), ("kt", C.c_double), ("gamman", C.c_double),
        ("gammat", C.c_double), ("xmu", C.c_double),
        ("normal_model", C.c_int), ("damping_model", C.c_int), ("tangential_model", C.c_int),
        ("g_kn", C.c_double), ("g_damp", C.c_double), ("g_kt", C.c_double),
        ("g_xt", C.c_double), ("g_mu", C.c_double),
    ]


class ServoParams(C.Structure):
    _fields_ = [
        ("enabled", C.c_int), ("nevery", C.c_int), ("couple", C.c_int), ("remap", C.c_int),
        ("normalize", C.c_int), ("max_rate", C.c_double),
        ("style", C.c_int * 3),
        ("rate", C.c_double * 3), ("vel", C.c_double * 3), ("scale", C.c_double * 3),
        ("flo", C.c_double * 3), ("fhi", C.c_double * 3), ("dlo", C.c_double * 3),
        ("dhi", C.c_double * 3), ("amp", C.c_double * 3), ("tper", C.c_double * 3),
        ("ptarget", C.c_double * 3), ("pgain", C.c_double * 3),
    ]


class WallParams(C.Structure):
    _fields_ = [
        ("axis", C.c_int), ("side", C.c_int), ("use_history", C.c_int),
        ("limit_damping", C.c_int), ("shear_axis", C.c_int),
        ("kn", C.c_double), ("kt", C.c_double), ("gamman", C.c_double),
        ("gammat", C.c_double), ("xmu", C.c_double),
        ("geometry", C.c_int), ("hertz", C.c_int), ("has_bounds", C.c_int), ("pad", C.c_int),
        ("c1", C.c_double), ("c2", C.c_double), ("radius", C.c_double),
        ("axis_lo", C.c_double), ("axis_hi", C.c_double),
    ]


class TlmParams(C.Structure):
    """orc_tlm_params (tlm.cpp)"""

    _fields_ = [("hertz", C.c_int), ("limit_damping", C.c_int),
                ("kn", C.c_double), ("kt", C.c_double), ("gamman", C.c_double),
                ("gammat", C.c_double), ("xmu", C.c_double),
                ("skin", C.c_double), ("dt", C.c_double), ("ptarget", C.c_double),
                ("pgain", C.c_double)]


STYLE_CODE = {"none": 0, "erate": 1, "trate": 2, "vel": 3, "wiggle": 4, "final": 5,
              "delta": 6, "scale": 7, "pressure": 8, "pressure/mean": 9}
COUPLE_CODE = {"none": 0, "xyz": 1, "xy": 2, "yz": 3, "xz": 4}
REMAP_CODE = {"x": 0, "v": 1, "none": 2}


def build() -> str:
    """Compile liboracle.so with its Makefile (idempotent)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_libs: dict = {}


def lib(threaded: bool = False):
    """The serial oracle, or with `threaded` its OpenMP build."""
    if threaded not in _libs:
        path = _OMP_LIB_PATH if threaded else _LIB_PATH
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_int, C.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_atoms.argtypes = [vp, dp, dp, dp, dp, dp, ip]
        L.orc_get_atoms.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.orc_set_forces.argtypes = [vp, dp, dp]
        L.orc_set_box.argtypes = [vp, dp, dp, ip]
        L.orc_set_h_rate.argtypes = [vp, dp, C.c_int]
        L.orc_get_box.argtypes = [vp, dp, dp, dp]
        L.orc_set_neigh.argtypes = [vp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_set_pair.argtypes = [vp, C.POINTER(PairParams)]
        L.orc_set_servo.argtypes = [vp, C.POINTER(ServoParams)]
        L.orc_set_cundall.argtypes = [vp, C.c_int, C.c_double, C.c_double]
        L.orc_run.argtypes = [vp, C.c_int, C.c_double, C.c_int]
        L.orc_init_simulation.argtypes = [vp, C.c_double]
        L.orc_step.argtypes = [vp, C.c_double]
        L.orc_servo_setup.argtypes = [vp, C.c_int, C.c_double]
        L.orc_nlist_build.argtypes = [vp]
        L.orc_nlist_decide.argtypes = [vp, C.c_int]
        L.orc_reneighbor.argtypes = [vp, C.c_int]
        L.orc_pbc.argtypes = [vp]
        L.orc_clear_forces.argtypes = [vp]
        L.orc_pair_compute.argtypes = [vp, C.c_double, C.c_int]
        L.orc_nve_initial.argtypes = [vp, C.c_double]
        L.orc_nve_final.argtypes = [vp, C.c_double]
        L.orc_cundall.argtypes = [vp]
        L.orc_virial.argtypes = [vp, C.c_int, C.c_int, dp]
        L.orc_servo_end_of_step.argtypes = [vp, C.c_double]
        L.orc_npairs.argtypes = [vp]
        L.orc_get_pairs.argtypes = [vp, ip, ip, ip, dp]
        L.orc_set_history.argtypes = [vp, ip, dp]
        L.orc_get_grid.argtypes = [vp, ip, dp, ip]
        L.orc_get_counters.argtypes = [vp, ip]
        L.orc_get_tensor.argtypes = [vp, dp]
        L.orc_num_threads.argtypes = []
        L.orc_contact_atom.argtypes = [vp, ip]
        L.orc_fabric.argtypes = [vp, dp, ip]
        L.orc_stress_atom.argtypes = [vp, C.c_int, dp, dp]
        L.orc_add_wall.argtypes = [vp, C.POINTER(WallParams), C.c_double, C.c_double,
                                   C.c_double, C.c_double]
        L.orc_wall_history.argtypes = [vp, C.c_int, dp, ip]
        L.orc_wall_post_force.argtypes = [vp, C.c_int, C.c_double, C.c_int]
        tp = C.POINTER(TlmParams)
        L.orc_tlm_forward.argtypes = [C.c_int, dp, dp, dp, dp, dp, ip, dp, dp, tp, C.c_int, dp,
                                      dp, dp, ip]
        L.orc_tlm_tangent.argtypes = [C.c_int, dp, dp, dp, dp, dp, ip, dp, dp, tp, C.c_int, dp,
                                      dp, dp, ip]
        for name in ("orc_det_sin", "orc_det_exp"):
            getattr(L, name).argtypes = [C.c_double]
            getattr(L, name).restype = C.c_double
        _libs[threaded] = L
    return _libs[threaded]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


# ------------------------------------------------------------ derivations

def mass_from_radius(radius, density):
    """atom.py:110-111: rmass = (4/3) pi r^3 rho, in the reference's numpy order."""
    rad = np.asarray(radius, dtype=np.float64)
    vol = (4.0 / 3.0) * np.pi * (rad**3)
    return vol * np.asarray(density, dtype=np.float64)


def legacy_pair(style, kn, gamman, kt=None, gammat=None, xmu=0.0, dampflag=1,
                limit_damping=False):
    """gran/hooke/history or gran/hertz/history, constructor order kn, gamman, kt, gammat.

    NULL (None) shorthands follow input.py:444-472: kt = 2/7 kn, gammat = gamman/2;
    dampflag 0 zeroes gammat (hooke_history.py:64-66).
    """
    if kt is None:
        kt = 2.0 / 7.0 * kn
    if gammat is None:
        gammat = 0.5 * gamman
    p = PairParams()
    p.style = {"hooke/history": 0, "hertz/history": 1}[style]
    p.limit_damping = 1 if limit_damping else 0
    p.kn, p.kt, p.gamman, p.xmu = float(kn), float(kt), float(gamman), float(xmu)
    p.gammat = 0.0 if int(dampflag) == 0 else float(gammat)
    return p


def granular_pair(normal, normal_coeffs, tangential="none", tangential_coeffs=None,
                  damping="viscoelastic", limit_damping=False):
    """PairGranular coefficient derivation, granular.py:94-269."""
    normals = {"hooke": 0, "hertz": 1, "hertz/material": 2}
    dampings = {"none": 0, "velocity": 1, "mass_velocity": 2, "viscoelastic": 3,
                "tsuji": 4, "coeff_restitution": 5}
    tangs = {"none": 0, "linear_nohistory": 1, "linear_history": 2, "mindlin": 3}
    p = PairParams()
    p.style = 2
    p.limit_damping = 1 if limit_damping else 0
    p.normal_model = normals[normal]
    p.damping_model = dampings[damping]
    p.tangential_model = tangs[tangential]
    emod = poiss = 0.0
    if normal == "hertz/material":
        emod, cor, poiss = (float(c) for c in normal_coeffs)
        mix_e = 1.0 / ((1 - poiss * poiss) / emod + (1 - poiss * poiss) / emod)
        p.g_kn = (4.0 / 3.0) * mix_e
    else:
        kn, cor = (float(c) for c in normal_coeffs)
        p.g_kn = kn
    if damping in ("tsuji", "coeff_restitution"):
        logcor = math.log(cor)
        if damping == "tsuji":
            d = (1.2728 - 4.2783 * cor + 11.087 * cor**2 - 22.348 * cor**3
                 + 27.467 * cor**4 - 18.022 * cor**5 + 4.8218 * cor**6)
            p.g_damp = d * math.sqrt(2.0)
        elif normal == "hooke":
            p.g_damp = -2 * logcor / math.sqrt(math.pi**2 + logcor**2)
        else:
            p.g_damp = (-1.22474487139158894067 * 1.82574185835055380345 * logcor
                        / math.sqrt(math.pi**2 + logcor**2))
    else:
        p.g_damp = cor
    tc = list(tangential_coeffs or [])
    if tangential == "linear_nohistory":
        p.g_xt, p.g_mu = float(tc[0]), float(tc[1])
    elif tangential in ("linear_history", "mindlin"):
        k_raw, xt, mu = tc
        p.g_xt, p.g_mu = float(xt), float(mu)
        if k_raw is None:
            mix_g = 1.0 / (2 * (2 - poiss) * (1 + poiss) / emod + 2 * (2 - poiss) * (1 + poiss) / emod)
            p.g_kt = 8.0 * mix_g
        else:
            p.g_kt = float(k_raw)
    return p


def servo_params(axes=None, couple="none", max_rate=0.0, normalize=False, remap="x",
                 nevery=1, p_target=None, p_gain=None):
    """FixDeformPressure settings for x/y/z (deform_pressure.py:151-238,332-356)."""
    s = ServoParams()
    s.enabled = 1
    s.nevery = int(nevery)
    s.remap = REMAP_CODE[remap]
    s.normalize = 1 if normalize else 0
    s.max_rate = float(max_rate)
    for i in range(3):
        s.scale[i] = 1.0
    if axes is None:
        target = 50000.0 if p_target is None else float(p_target)
        gain = 1e-4 if p_gain is None else float(p_gain)
        for i in range(3):
            s.style[i] = STYLE_CODE["pressure"]
            s.ptarget[i] = target
            s.pgain[i] = gain
        if couple == "none":
            couple = "xyz"
    else:
        names = {"x": 0, "y": 1, "z": 2}
        for name, spec in axes.items():
            i = names[name]
            s.style[i] = STYLE_CODE[spec["style"]]
            for key in ("rate", "vel", "scale", "flo", "fhi", "dlo", "dhi", "ptarget", "pgain"):
                if key in spec:
                    getattr(s, key)[i] = float(spec[key])
            if "amplitude" in spec:
                s.amp[i] = float(spec["amplitude"])
            if "tperiod" in spec:
                s.tper[i] = float(spec["tperiod"])
        if couple != "none":
            idx = [i for i, c in enumerate("xyz") if c in couple]
            ref = next(i for i in idx if s.style[i] in (8, 9))
            for i in idx:
                if s.style[i] == 0:
                    s.style[i] = s.style[ref]
                    s.ptarget[i] = s.ptarget[ref]
                    s.pgain[i] = s.pgain[ref]
    s.couple = COUPLE_CODE[couple]
    return s


# ------------------------------------------------------------------ handle

class OracleSim:
    """One simulated system held by the C++ oracle."""

    def __init__(self, n: int, max_nb: int = 64, threaded: bool = False):
        self.n = int(n)
        self.max_nb = int(max_nb)
        self._L = lib(threaded)
        self._h = self._L.orc_create(self.n, self.max_nb)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.orc_destroy(h)
            self._h = None

    # state -----------------------------------------------------------------
    def set_atoms(self, x, v, omega, radius, rmass, tag=None):
        n = self.n
        arr = lambda a, s: np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(s))
        xx, vv, oo = arr(x, (n, 3)), arr(v, (n, 3)), arr(omega, (n, 3))
        rr, mm = arr(radius, (n,)), arr(rmass, (n,))
        tt = np.ascontiguousarray(
            np.arange(1, n + 1, dtype=np.int32) if tag is None else np.asarray(tag, dtype=np.int32))
        self._L.orc_set_atoms(self._h, _dp(xx), _dp(vv), _dp(oo), _dp(rr), _dp(mm), _ip(tt))

    def set_forces(self, f, torque):
        ff = np.ascontiguousarray(np.asarray(f, dtype=np.float64))
        tt = np.ascontiguousarray(np.asarray(torque, dtype=np.float64))
        self._L.orc_set_forces(self._h, _dp(ff), _dp(tt))

    def atoms(self) -> dict:
        n = self.n
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "omega", "f", "torque")}
        out["virial"] = np.zeros((n, 6))
        self._L.orc_get_atoms(self._h, *(_dp(out[k]) for k in ("x", "v", "omega", "f", "torque", "virial")))
        return out

    def set_box(self, lo, hi, periodic=(1, 1, 1)):
        lo_ = np.asarray(lo, dtype=np.float64)
        hi_ = np.asarray(hi, dtype=np.float64)
        per = np.asarray(periodic, dtype=np.int32)
        self._L.orc_set_box(self._h, _dp(lo_), _dp(hi_), _ip(per))

    def set_h_rate(self, h_rate, vremap):
        h = np.asarray(h_rate, dtype=np.float64)
        self._L.orc_set_h_rate(self._h, _dp(h), int(vremap))

    def box(self):
        lo, hi, hr = np.zeros(3), np.zeros(3), np.zeros(6)
        self._L.orc_get_box(self._h, _dp(lo), _dp(hi), _dp(hr))
        return lo, hi, hr

    def set_neigh(self, skin, delay=0, every=1, check=True, max_grid_cells=200000):
        self._L.orc_set_neigh(self._h, float(skin), int(delay), int(every), 1 if check else 0,
                            int(max_grid_cells))

    def set_pair(self, p: PairParams):
        self._pair = p
        self._L.orc_set_pair(self._h, C.byref(p))

    def set_servo(self, s: ServoParams):
        self._servo = s
        self._L.orc_set_servo(self._h, C.byref(s))

    def set_cundall(self, gamma_lin, gamma_ang, enabled=True):
        self._L.orc_set_cundall(self._h, 1 if enabled else 0, float(gamma_lin), float(gamma_ang))

    # stages ----------------------------------------------------------------
    def run(self, nsteps, dt, run_pre=True) -> int:
        return self._L.orc_run(self._h, int(nsteps), float(dt), 1 if run_pre else 0)

    def init_simulation(self, dt) -> int:
        return self._L.orc_init_simulation(self._h, float(dt))

    def step(self, dt) -> int:
        return self._L.orc_step(self._h, float(dt))

    def servo_setup(self, nsteps_total, dt):
        self._L.orc_servo_setup(self._h, int(nsteps_total), float(dt))

    def nlist_build(self) -> int:
        return self._L.orc_nlist_build(self._h)

    def decide(self, timestep) -> bool:
        return bool(self._L.orc_nlist_decide(self._h, int(timestep)))

    def reneighbor(self, force=False) -> int:
        return self._L.orc_reneighbor(self._h, 1 if force else 0)

    def pbc(self):
        self._L.orc_pbc(self._h)

    def clear_forces(self):
        self._L.orc_clear_forces(self._h)

    def pair_compute(self, dt, shearupdate=True):
        self._L.orc_pair_compute(self._h, float(dt), 1 if shearupdate else 0)

    def nve_initial(self, dt):
        self._L.orc_nve_initial(self._h, float(dt))

    def nve_final(self, dt):
        self._L.orc_nve_final(self._h, float(dt))

    def cundall(self):
        self._L.orc_cundall(self._h)

    def virial(self, keflag=True, mode=0):
        out = np.zeros(6)
        self._L.orc_virial(self._h, 1 if keflag else 0, int(mode), _dp(out))
        return out

    def servo_end_of_step(self, dt) -> int:
        return self._L.orc_servo_end_of_step(self._h, float(dt))

    # fix wall/gran ----------------------------------------------------------
    def add_wall(self, axis, side, coord, kn, gamman, kt, gammat, xmu, dampflag=1,
                 history=True, limit_damping=False, shear_axis=-1, shear_vel=0.0,
                 wiggle_amplitude=0.0, wiggle_period=0.0) -> int:
        """FixWallGran (wall_gran.py:57-111), applied after damping/cundall in
        every post_force; returns the wall's index."""
        p = WallParams()
        p.axis, p.side = int(axis), int(side)
        p.use_history, p.limit_damping = (1 if history else 0), (1 if limit_damping else 0)
        p.shear_axis = int(shear_axis)
        p.kn, p.kt, p.gamman, p.xmu = float(kn), float(kt), float(gamman), float(xmu)
        p.gammat = 0.0 if int(dampflag) == 0 else float(gammat)
        return self._L.orc_add_wall(self._h, C.byref(p), float(coord), float(shear_vel),
                                  float(wiggle_amplitude), float(wiggle_period))

    def add_wall_cylinder(self, axis, c1, c2, radius, side, kn, gamman, kt=None, gammat=None,
                          xmu=0.0, dampflag=1, fstyle="hooke", history=None,
                          limit_damping=False, axis_lo=None, axis_hi=None) -> int:
        """FixWallGranRegion (wall_gran_region.py:45-112): side -1 inside, +1
        outside; kt NULL = 2/7 kn, gammat NULL = gamman / 2."""
        p = WallParams()
        p.geometry, p.axis = 1, int(axis)
        p.side = -1 if side in ("in", -1) else 1
        p.c1, p.c2, p.radius = float(c1), float(c2), float(radius)
        p.has_bounds = 1 if (axis_lo is not None or axis_hi is not None) else 0
        p.axis_lo = float(axis_lo) if axis_lo is not None else -1e30
        p.axis_hi = float(axis_hi) if axis_hi is not None else 1e30
        p.hertz = 1 if "hertz" in fstyle else 0
        p.kn, p.gamman, p.xmu = float(kn), float(gamman), float(xmu)
        p.kt = float(kt) if kt is not None else 2.0 / 7.0 * float(kn)
        gt = float(gammat) if gammat is not None else 0.5 * float(gamman)
        p.gammat = 0.0 if int(dampflag) == 0 else gt
        p.use_history = 1 if (fstyle.endswith("history") if history is None else history) else 0
        p.limit_damping = 1 if limit_damping else 0
        p.shear_axis = -1
        return self._L.orc_add_wall(self._h, C.byref(p), 0.0, 0.0, 0.0, 0.0)

    def wall_history(self, idx):
        sh = np.zeros((self.n, 3))
        tc = np.zeros(self.n, dtype=np.int32)
        self._L.orc_wall_history(self._h, int(idx), _dp(sh), _ip(tc))
        return sh, tc

    def wall_post_force(self, idx, dt, history_update=True):
        self._L.orc_wall_post_force(self._h, int(idx), float(dt), 1 if history_update else 0)

    # per-contact computes ----------------------------------------------------
    def contact_atom(self):
        """ComputeContactAtom.compute (contact_atom.py:19-64): int32[n]."""
        out = np.zeros(self.n, dtype=np.int32)
        self._L.orc_contact_atom(self._h, _ip(out))
        return out

    def fabric(self):
        """ComputeFabric.compute (fabric.py:26-97): (phi[6], Nc)."""
        phi = np.zeros(6)
        nc = np.zeros(1, dtype=np.int32)
        self._L.orc_fabric(self._h, _dp(phi), _ip(nc))
        return phi, int(nc[0])

    def stress_atom(self, keflag=False):
        """ComputeStressAtom.compute (stress_atom.py:30-90): (stress[n,6], sum[6])."""
        st = np.zeros((self.n, 6))
        tot = np.zeros(6)
        self._L.orc_stress_atom(self._h, 1 if keflag else 0, _dp(st), _dp(tot))
        return st, tot

    # inspection ------------------------------------------------------------
    def pairs(self):
        m = self._L.orc_npairs(self._h)
        pi, pj, pa = (np.zeros(m, dtype=np.int32) for _ in range(3))
        sh = np.zeros((m, 3))
        self._L.orc_get_pairs(self._h, _ip(pi), _ip(pj), _ip(pa), _dp(sh))
        return pi, pj, pa, sh

    def set_history(self, partner, shear):
        pa = np.ascontiguousarray(np.asarray(partner, dtype=np.int32))
        sh = np.ascontiguousarray(np.asarray(shear, dtype=np.float64))
        self._L.orc_set_history(self._h, _ip(pa), _dp(sh))

    def grid(self):
        g = np.zeros(3, dtype=np.int32)
        cs = np.zeros(3)
        cell_of = np.zeros(self.n, dtype=np.int32)
        self._L.orc_get_grid(self._h, _ip(g), _dp(cs), _ip(cell_of))
        return g, cs, cell_of

    def counters(self) -> dict:
        c = np.zeros(6, dtype=np.int32)
        self._L.orc_get_counters(self._h, _ip(c))
        keys = ("timestep", "build_count", "last_build_step", "npairs", "rebuilds_in_run", "has_built")
        return {k: int(v) for k, v in zip(keys, c)}

    def last_pressure_tensor(self):
        out = np.zeros(6)
        self._L.orc_get_tensor(self._h, _dp(out))
        return out


def det_sin(x: float) -> float:
    """The servo's sin (wiggle), evaluated by the fixed method the device uses."""
    return lib().orc_det_sin(float(x))


def det_exp(x: float) -> float:
    """The servo's exp (trate), evaluated by the fixed method the device uses."""
    return lib().orc_det_exp(float(x))


# ------------------------------------------------ tangent-linear run (tlm.cpp)

def _tlm_args(w, coeffs, style, skin, dt, p_target, p_gain):
    n = len(w["x"])
    arr = lambda a, shape: np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(shape))
    r = arr(np.broadcast_to(np.asarray(w["radius"], dtype=np.float64), (n,)), (n,))
    m = arr(mass_from_radius(r, w["density"]), (n,))
    tag = np.ascontiguousarray(np.asarray(w["tag"], dtype=np.int32))
    kn, kt, gn, gt, xmu = (float(c) for c in coeffs)
    p = TlmParams(1 if style == "hertz/history" else 0, 0, kn, kt, gn, gt, xmu, float(skin),
                  float(dt), float(p_target), float(p_gain))
    return (n, arr(w["x"], (n, 3)), arr(w["v"], (n, 3)), arr(w["omega"], (n, 3)), r, m, tag,
            arr(w["boxlo"], (3,)), arr(w["boxhi"], (3,)), p)


def tlm_forward(w, coeffs, nsteps, style="hertz/history", skin=0.1, dt=1e-4, p_target=1.0,
                p_gain=1e-4):
    """Simulation.run(nsteps) of the isotropic compression by the
    tangent-linear restatement's FP64 form: (x, v, pressure tensor, builds)."""
    n, x, v, om, r, m, tag, lo, hi, p = _tlm_args(w, coeffs, style, skin, dt, p_target, p_gain)
    xo, vo, t = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(6)
    b = np.zeros(1, dtype=np.int32)
    rc = lib().orc_tlm_forward(n, _dp(x), _dp(v), _dp(om), _dp(r), _dp(m), _ip(tag), _dp(lo),
                               _dp(hi), C.byref(p), int(nsteps), _dp(xo), _dp(vo), _dp(t), _ip(b))
    if rc:
        raise RuntimeError("the box collapsed")
    return xo, vo, t, int(b[0])


def tlm_tangent(w, coeffs, nsteps, dradius=None, style="hertz/history", skin=0.1, dt=1e-4,
                p_target=1.0, p_gain=1e-4):
    """The same run differentiated forward: (pressure tensor [6], d tensor /
    d(kn, kt, gamman, gammat, xmu, radius direction) [6 x 6], builds)."""
    n, x, v, om, r, m, tag, lo, hi, p = _tlm_args(w, coeffs, style, skin, dt, p_target, p_gain)
    t, dt6 = np.zeros(6), np.zeros((6, 6))
    b = np.zeros(1, dtype=np.int32)
    dr = None if dradius is None else np.ascontiguousarray(np.asarray(dradius, dtype=np.float64))
    rc = lib().orc_tlm_tangent(n, _dp(x), _dp(v), _dp(om), _dp(r), _dp(m), _ip(tag), _dp(lo),
                               _dp(hi), C.byref(p), int(nsteps),
                               _dp(dr) if dr is not None else None, _dp(t), _dp(dt6), _ip(b))
    if rc:
        raise RuntimeError("the box collapsed")
    return t, dt6, int(b[0])


def num_threads(threaded: bool = True) -> int:
    """Host threads the (threaded) oracle's loops run on."""
    return int(lib(threaded).orc_num_threads())


def from_workload(w: dict, max_nb: int = 64, threaded: bool = False) -> OracleSim:
    """An OracleSim loaded with a synthetic workload dict (see taichimps_amd.workloads)."""
    o = OracleSim(len(w["x"]), max_nb=max_nb, threaded=threaded)
    o.set_atoms(w["x"], w["v"], w["omega"], w["radius"],
                mass_from_radius(w["radius"], w["density"]), w["tag"])
    o.set_box(w["boxlo"], w["boxhi"], (1, 1, 1))
    return o
