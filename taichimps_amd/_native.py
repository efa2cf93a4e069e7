"""ctypes binding of libtmps_hip.so (the C ABI declared in include/tmps.h).

The library owns all device memory; these wrappers hand it borrowed numpy
buffers. Status codes become the exceptions the reference raises for the same
conditions (ValueError for configuration, RuntimeError at run time).

torch is imported before the library is loaded when it is importable, so the
process holds exactly one HIP runtime (torch ships its own copy under the same
soname); the library itself does not use torch.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtmps_hip.so")

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_OVERFLOW = -3
ERR_BOX = -4
ERR_CAPACITY = -5
ERR_UNSUPPORTED = -6

FP64 = 0
FP32_SPLIT = 1

F_X, F_V, F_OMEGA, F_F, F_TORQUE, F_RADIUS, F_RMASS, F_TAG, F_VIRIAL, F_X0 = range(10)
FIELD_WIDTH = {F_X: 3, F_V: 3, F_OMEGA: 3, F_F: 3, F_TORQUE: 3, F_RADIUS: 1, F_RMASS: 1,
               F_TAG: 1, F_VIRIAL: 6, F_X0: 3}

PAIR_HOOKE_HISTORY = 0
PAIR_HERTZ_HISTORY = 1
PAIR_GRANULAR = 2

# the exported entry points (include/tmps.h)
SYMBOLS = (
    "tmps_version", "tmps_last_error", "tmps_device_name", "tmps_create", "tmps_destroy",
    "tmps_set_nlocal", "tmps_set_field", "tmps_get_field", "tmps_set_field_device",
    "tmps_get_field_device", "tmps_sync", "tmps_set_box",
    "tmps_get_box", "tmps_pbc", "tmps_clear_forces", "tmps_neigh_config", "tmps_neigh_check",
    "tmps_neigh_decide", "tmps_neigh_build", "tmps_neigh_info", "tmps_neigh_get_pairs",
    "tmps_neigh_get_counts", "tmps_neigh_get_bins", "tmps_history_reset", "tmps_history_save",
    "tmps_history_restore", "tmps_history_set_pairs", "tmps_pair_set", "tmps_pair_compute",
    "tmps_nve_initial", "tmps_nve_final", "tmps_cundall", "tmps_servo_set", "tmps_servo_setup",
    "tmps_servo_end_of_step", "tmps_virial_reduce", "tmps_ke", "tmps_run", "tmps_get_counters",
    "tmps_set_counters", "tmps_last_pressure", "tmps_set_profiling", "tmps_kernel_times",
    "tmps_set_launch_mode", "tmps_adjoint_forward", "tmps_adjoint_backward",
    "tmps_batch_create", "tmps_batch_destroy", "tmps_batch_run", "tmps_batch_tuning",
    "tmps_contact_atom", "tmps_fabric", "tmps_stress_atom", "tmps_history_seed",
    "tmps_wall_create", "tmps_wall_destroy", "tmps_wall_reset", "tmps_wall_post_force",
    "tmps_wall_get_history", "tmps_prepare_run", "tmps_set_stream", "tmps_get_stream",
    "tmps_fabric_noalloc", "tmps_run_set_walls",
)


class PairParams(C.Structure):
    """tmps_pair_params"""

    _fields_ = [
        ("style", C.c_int), ("limit_damping", C.c_int),
        ("kn", C.c_double), ("kt", C.c_double), ("gamman", C.c_double),
        ("gammat", C.c_double), ("xmu", C.c_double),
        ("normal_model", C.c_int), ("damping_model", C.c_int), ("tangential_model", C.c_int),
        ("g_kn", C.c_double), ("g_damp", C.c_double), ("g_kt", C.c_double),
        ("g_xt", C.c_double), ("g_mu", C.c_double),
    ]


class ServoParams(C.Structure):
    """tmps_servo_params"""

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
    """tmps_wall_params"""

    _fields_ = [
        ("axis", C.c_int), ("side", C.c_int), ("use_history", C.c_int),
        ("limit_damping", C.c_int), ("shear_axis", C.c_int),
        ("kn", C.c_double), ("kt", C.c_double), ("gamman", C.c_double),
        ("gammat", C.c_double), ("xmu", C.c_double),
        ("geometry", C.c_int), ("hertz", C.c_int), ("has_bounds", C.c_int), ("pad", C.c_int),
        ("c1", C.c_double), ("c2", C.c_double), ("radius", C.c_double),
        ("axis_lo", C.c_double), ("axis_hi", C.c_double),
    ]


class Counters(C.Structure):
    """tmps_counters"""

    _fields_ = [(name, C.c_int) for name in (
        "timestep", "has_built", "build_count", "last_build_step", "current_step", "setup_done",
        "npairs", "nslots", "ntouch", "rebuilds", "servo_step_count", "servo_nsteps")]


_lib = None


def _preload_torch():
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def load(path: str | None = None):
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: the HIP library has not been built "
            "(run __graft_entry__.build() or make -C taichimps_amd/csrc)")
    _preload_torch()
    L = C.CDLL(p)
    vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
    sig = {
        "tmps_version": (C.c_char_p, []),
        "tmps_last_error": (C.c_char_p, []),
        "tmps_device_name": (C.c_int, [C.c_char_p, C.c_int]),
        "tmps_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
        "tmps_destroy": (C.c_int, [vp]),
        "tmps_set_nlocal": (C.c_int, [vp, C.c_int]),
        "tmps_set_field": (C.c_int, [vp, C.c_int, C.c_void_p, C.c_int]),
        "tmps_get_field": (C.c_int, [vp, C.c_int, C.c_void_p, C.c_int]),
        "tmps_set_field_device": (C.c_int, [vp, C.c_int, C.c_void_p, C.c_int]),
        "tmps_get_field_device": (C.c_int, [vp, C.c_int, C.c_void_p, C.c_int]),
        "tmps_sync": (C.c_int, [vp]),
        "tmps_set_box": (C.c_int, [vp, dp, dp, ip, dp, C.c_int]),
        "tmps_get_box": (C.c_int, [vp, dp, dp, dp, ip]),
        "tmps_pbc": (C.c_int, [vp]),
        "tmps_clear_forces": (C.c_int, [vp]),
        "tmps_neigh_config": (C.c_int, [vp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "tmps_neigh_check": (C.c_int, [vp, ip]),
        "tmps_neigh_decide": (C.c_int, [vp, C.c_int, C.c_int, ip]),
        "tmps_neigh_build": (C.c_int, [vp]),
        "tmps_neigh_info": (C.c_int, [vp, ip, dp, ip]),
        "tmps_neigh_get_pairs": (C.c_int, [vp, ip, ip, ip, dp, C.c_int]),
        "tmps_neigh_get_counts": (C.c_int, [vp, ip, C.c_int]),
        "tmps_neigh_get_bins": (C.c_int, [vp, ip, ip, ip]),
        "tmps_history_reset": (C.c_int, [vp]),
        "tmps_history_save": (C.c_int, [vp]),
        "tmps_history_restore": (C.c_int, [vp]),
        "tmps_history_set_pairs": (C.c_int, [vp, ip, dp, C.c_int]),
        "tmps_pair_set": (C.c_int, [vp, C.POINTER(PairParams)]),
        "tmps_pair_compute": (C.c_int, [vp, C.c_double, C.c_int, C.c_int]),
        "tmps_nve_initial": (C.c_int, [vp, C.c_double]),
        "tmps_nve_final": (C.c_int, [vp, C.c_double]),
        "tmps_cundall": (C.c_int, [vp, C.c_double, C.c_double]),
        "tmps_servo_set": (C.c_int, [vp, C.POINTER(ServoParams)]),
        "tmps_servo_setup": (C.c_int, [vp, C.c_int, C.c_double]),
        "tmps_servo_end_of_step": (C.c_int, [vp, C.c_double, dp]),
        "tmps_virial_reduce": (C.c_int, [vp, C.c_int, dp]),
        "tmps_contact_atom": (C.c_int, [vp, ip, C.POINTER(C.c_longlong)]),
        "tmps_history_seed": (C.c_int, [vp, C.c_int, ip, ip, dp]),
        "tmps_wall_create": (C.c_int, [vp, C.POINTER(WallParams), C.POINTER(vp)]),
        "tmps_wall_destroy": (C.c_int, [vp]),
        "tmps_wall_reset": (C.c_int, [vp]),
        "tmps_wall_post_force": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_int]),
        "tmps_wall_get_history": (C.c_int, [vp, dp, ip, C.c_int]),
        "tmps_run_set_walls": (C.c_int, [vp, C.c_int, C.POINTER(vp), ip, dp, dp, C.c_double]),
        "tmps_fabric": (C.c_int, [vp, dp, ip]),
        "tmps_fabric_noalloc": (C.c_int, [vp, dp, ip]),
        "tmps_stress_atom": (C.c_int, [vp, C.c_int, dp, dp]),
        "tmps_ke": (C.c_int, [vp, dp, dp]),
        "tmps_run": (C.c_int, [vp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double]),
        "tmps_prepare_run": (C.c_int, [vp, C.c_int, C.c_double, C.c_int, C.c_double,
                                       C.c_double]),
        "tmps_set_stream": (C.c_int, [vp, vp]),
        "tmps_get_stream": (C.c_int, [vp, C.POINTER(vp)]),
        "tmps_get_counters": (C.c_int, [vp, C.POINTER(Counters)]),
        "tmps_set_counters": (C.c_int, [vp, C.POINTER(Counters)]),
        "tmps_last_pressure": (C.c_int, [vp, dp]),
        "tmps_set_profiling": (C.c_int, [vp, C.c_int]),
        "tmps_kernel_times": (C.c_int, [vp, dp, ip]),
        "tmps_set_launch_mode": (C.c_int, [vp, C.c_int]),
        "tmps_adjoint_forward": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, dp, dp]),
        "tmps_adjoint_backward": (C.c_int, [vp, dp, dp, dp, dp, dp, dp]),
        "tmps_batch_create": (C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(vp)]),
        "tmps_batch_destroy": (C.c_int, [vp]),
        "tmps_batch_run": (C.c_int, [vp, C.c_int, C.c_double, C.c_int, C.c_int]),
        "tmps_batch_tuning": (C.c_int, [vp, ip, ip, dp, ip, ip, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error() -> str:
    msg = load().tmps_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> None:
    """Raise the reference's exception type for a non-zero status."""
    if rc == OK:
        return
    msg = last_error()
    if rc in (ERR_ARG, ERR_CAPACITY):
        raise ValueError(msg)
    if rc == ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class NativeSystem:
    """One tmps_system: the device-resident state of one particle system."""

    def __init__(self, max_atoms: int, max_neighbors_per_atom: int, precision: int):
        L = load()
        h = C.c_void_p()
        check(L.tmps_create(int(max_atoms), int(max_neighbors_per_atom), int(precision), C.byref(h)))
        self._h = h
        self._L = L
        self.max_atoms = int(max_atoms)
        self.precision = int(precision)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._L.tmps_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    # fields -----------------------------------------------------------------
    def set_nlocal(self, n):
        check(self._L.tmps_set_nlocal(self._h, int(n)))

    def set_field(self, field, values, count):
        width = FIELD_WIDTH[field]
        dtype = np.int32 if field == F_TAG else np.float64
        a = np.ascontiguousarray(np.asarray(values, dtype=dtype))
        if a.size < count * width:
            raise ValueError("field data shorter than count")
        check(self._L.tmps_set_field(self._h, field, a.ctypes.data_as(C.c_void_p), int(count)))

    def get_field(self, field, count):
        width = FIELD_WIDTH[field]
        dtype = np.int32 if field == F_TAG else np.float64
        shape = (count,) if width == 1 else (count, width)
        out = np.zeros(shape, dtype=dtype)
        check(self._L.tmps_get_field(self._h, field, out.ctypes.data_as(C.c_void_p), int(count)))
        return out

    def set_field_device(self, field, tensor, count, sync=True):
        """tmps_set_field_device from a contiguous GPU torch tensor (float64;
        int32 for tags), borrowed for the call. With `sync` the work queued
        on torch's current stream is finished first and the system's copy
        after, so the tensor is free on return; without it the caller orders
        them (e.g. one stream for both, set_stream)."""
        import torch
        width = FIELD_WIDTH[field]
        dtype = torch.int32 if field == F_TAG else torch.float64
        if not tensor.is_cuda or tensor.dtype != dtype or not tensor.is_contiguous():
            raise ValueError(f"field data must be a contiguous {dtype} tensor on the GPU")
        if tensor.numel() < count * width:
            raise ValueError("field data shorter than count")
        if sync:
            torch.cuda.current_stream(tensor.device).synchronize()
        check(self._L.tmps_set_field_device(self._h, field, C.c_void_p(tensor.data_ptr()),
                                            int(count)))
        if sync:
            self.sync()

    def get_field_device(self, field, count, sync=True):
        """tmps_get_field_device into a new GPU torch tensor (float64; int32
        for tags), shaped as get_field's array. With `sync` the system's
        stream is finished before the tensor is returned."""
        import torch
        width = FIELD_WIDTH[field]
        dtype = torch.int32 if field == F_TAG else torch.float64
        shape = (count,) if width == 1 else (count, width)
        out = torch.empty(shape, dtype=dtype, device="cuda")
        if sync:
            torch.cuda.current_stream(out.device).synchronize()
        check(self._L.tmps_get_field_device(self._h, field, C.c_void_p(out.data_ptr()),
                                            int(count)))
        if sync:
            self.sync()
        return out

    def sync(self):
        check(self._L.tmps_sync(self._h))

    # box --------------------------------------------------------------------
    def set_box(self, lo, hi, periodic, h_rate=None, vremap=False):
        lo_ = np.ascontiguousarray(lo, dtype=np.float64)
        hi_ = np.ascontiguousarray(hi, dtype=np.float64)
        per = np.ascontiguousarray(periodic, dtype=np.int32)
        hr = np.ascontiguousarray(np.zeros(6) if h_rate is None else h_rate, dtype=np.float64)
        check(self._L.tmps_set_box(self._h, _dp(lo_), _dp(hi_), _ip(per), _dp(hr),
                                   1 if vremap else 0))

    def get_box(self):
        lo, hi, hr = np.zeros(3), np.zeros(3), np.zeros(6)
        vr = np.zeros(1, dtype=np.int32)
        check(self._L.tmps_get_box(self._h, _dp(lo), _dp(hi), _dp(hr), _ip(vr)))
        return lo, hi, hr, bool(vr[0])

    def pbc(self):
        check(self._L.tmps_pbc(self._h))

    def clear_forces(self):
        check(self._L.tmps_clear_forces(self._h))

    # neighbor list ----------------------------------------------------------
    def neigh_config(self, skin, delay, every, check_, max_grid_cells, max_nb):
        check(self._L.tmps_neigh_config(self._h, float(skin), int(delay), int(every),
                                        1 if check_ else 0, int(max_grid_cells), int(max_nb)))

    def neigh_check(self) -> bool:
        m = np.zeros(1, dtype=np.int32)
        check(self._L.tmps_neigh_check(self._h, _ip(m)))
        return bool(m[0])

    def neigh_decide(self, timestep) -> bool:
        r = np.zeros(1, dtype=np.int32)
        have = timestep is not None
        check(self._L.tmps_neigh_decide(self._h, int(timestep or 0), 1 if have else 0, _ip(r)))
        return bool(r[0])

    def neigh_build(self):
        check(self._L.tmps_neigh_build(self._h))

    def neigh_info(self):
        g = np.zeros(3, dtype=np.int32)
        cs = np.zeros(3)
        npairs = np.zeros(1, dtype=np.int32)
        check(self._L.tmps_neigh_info(self._h, _ip(g), _dp(cs), _ip(npairs)))
        return g, cs, int(npairs[0])

    def neigh_pairs(self):
        _, _, m = self.neigh_info()
        pi, pj, pa = (np.zeros(m, dtype=np.int32) for _ in range(3))
        sh = np.zeros((m, 3))
        if m:
            check(self._L.tmps_neigh_get_pairs(self._h, _ip(pi), _ip(pj), _ip(pa), _dp(sh), m))
        return pi, pj, pa, sh

    def neigh_counts(self, n):
        out = np.zeros(n, dtype=np.int32)
        if n:
            check(self._L.tmps_neigh_get_counts(self._h, _ip(out), int(n)))
        return out

    def neigh_bins(self, n):
        g, _, _ = self.neigh_info()
        ncell = int(g[0] * g[1] * g[2])
        cell_of = np.zeros(n, dtype=np.int32)
        start = np.zeros(ncell + 1, dtype=np.int32)
        parts = np.zeros(n, dtype=np.int32)
        check(self._L.tmps_neigh_get_bins(self._h, _ip(cell_of), _ip(start), _ip(parts)))
        return cell_of, start, parts

    # history ----------------------------------------------------------------
    def history_reset(self):
        check(self._L.tmps_history_reset(self._h))

    def history_save(self):
        check(self._L.tmps_history_save(self._h))

    def history_restore(self):
        check(self._L.tmps_history_restore(self._h))

    def history_set_pairs(self, partner, shear):
        pa = np.ascontiguousarray(partner, dtype=np.int32)
        sh = np.ascontiguousarray(shear, dtype=np.float64)
        check(self._L.tmps_history_set_pairs(self._h, _ip(pa), _dp(sh), int(len(pa))))

    # stages -----------------------------------------------------------------
    def history_seed(self, start, nbr, shear):
        """Seed the saved history: CSR rows of partner indices and their shear."""
        st = np.ascontiguousarray(start, dtype=np.int32)
        nb = np.ascontiguousarray(nbr, dtype=np.int32)
        sh = np.ascontiguousarray(shear, dtype=np.float64).reshape(-1, 3)
        n = len(st) - 1
        nb_ = nb if len(nb) else np.zeros(1, dtype=np.int32)
        sh_ = sh if len(sh) else np.zeros((1, 3))
        check(self._L.tmps_history_seed(self._h, n, _ip(st), _ip(nb_), _dp(sh_)))

    def pair_set(self, p: PairParams):
        check(self._L.tmps_pair_set(self._h, C.byref(p)))

    def pair_compute(self, dt, shearupdate=True, accumulate=True):
        check(self._L.tmps_pair_compute(self._h, float(dt), 1 if shearupdate else 0,
                                        1 if accumulate else 0))

    def nve_initial(self, dt):
        check(self._L.tmps_nve_initial(self._h, float(dt)))

    def nve_final(self, dt):
        check(self._L.tmps_nve_final(self._h, float(dt)))

    def cundall(self, gl, ga):
        check(self._L.tmps_cundall(self._h, float(gl), float(ga)))

    def servo_set(self, p: ServoParams):
        check(self._L.tmps_servo_set(self._h, C.byref(p)))

    def servo_setup(self, nsteps_total, dt):
        check(self._L.tmps_servo_setup(self._h, int(nsteps_total), float(dt)))

    def servo_end_of_step(self, dt):
        t = np.zeros(6)
        check(self._L.tmps_servo_end_of_step(self._h, float(dt), _dp(t)))
        return t

    def virial(self, keflag=True):
        out = np.zeros(6)
        check(self._L.tmps_virial_reduce(self._h, 1 if keflag else 0, _dp(out)))
        return out

    def ke(self):
        a, b = np.zeros(1), np.zeros(1)
        check(self._L.tmps_ke(self._h, _dp(a), _dp(b)))
        return float(a[0]), float(b[0])

    def contact_atom(self, n):
        """Per-atom overlapping partners (int32[n]) and their total."""
        out = np.zeros(max(n, 1), dtype=np.int32)
        total = C.c_longlong(0)
        check(self._L.tmps_contact_atom(self._h, _ip(out), C.byref(total)))
        return out[:n], int(total.value)

    def fabric(self):
        """Contact-normal fabric [xx yy zz xy xz yz] and the contact count."""
        phi = np.zeros(6)
        nc = np.zeros(1, dtype=np.int32)
        check(self._L.tmps_fabric(self._h, _dp(phi), _ip(nc)))
        return phi, int(nc[0])

    def stress_atom(self, n, keflag=False, per_atom=True):
        """Per-atom stress (n x 6, or None) and its sum over atoms."""
        st = np.zeros((max(n, 1), 6)) if per_atom else None
        tot = np.zeros(6)
        check(self._L.tmps_stress_atom(self._h, 1 if keflag else 0,
                                       _dp(st) if per_atom else None, _dp(tot)))
        return (st[:n] if per_atom else None), tot

    def run(self, nsteps, dt, run_pre=True, cundall=None):
        gl, ga = (0.0, 0.0) if cundall is None else cundall
        check(self._L.tmps_run(self._h, int(nsteps), float(dt), 1 if run_pre else 0,
                               0 if cundall is None else 1, float(gl), float(ga)))

    def set_run_walls(self, walls, dt):
        """The walls of the next device-resident runs: (NativeWall,
        before_cundall, coord, vwall_s) in the fixes' order; [] clears them."""
        n = len(walls)
        hs = (C.c_void_p * max(n, 1))(*[w[0]._h.value for w in walls])
        bc = np.array([1 if w[1] else 0 for w in walls] or [0], dtype=np.int32)
        co = np.array([float(w[2]) for w in walls] or [0.0])
        vs = np.array([float(w[3]) for w in walls] or [0.0])
        check(self._L.tmps_run_set_walls(self._h, n, hs, _ip(bc), _dp(co), _dp(vs), float(dt)))

    def prepare_run(self, nsteps, dt, cundall=None):
        gl, ga = (0.0, 0.0) if cundall is None else cundall
        check(self._L.tmps_prepare_run(self._h, int(nsteps), float(dt),
                                       0 if cundall is None else 1, float(gl), float(ga)))

    def set_stream(self, handle):
        """Run on a caller's HIP stream (an integer handle, e.g. a torch
        stream's cuda_stream), or on the system's own with None."""
        check(self._L.tmps_set_stream(self._h, C.c_void_p(handle) if handle else None))

    def stream(self) -> int:
        out = C.c_void_p()
        check(self._L.tmps_get_stream(self._h, C.byref(out)))
        return int(out.value or 0)

    def counters(self) -> dict:
        c = Counters()
        check(self._L.tmps_get_counters(self._h, C.byref(c)))
        return {name: getattr(c, name) for name, _ in Counters._fields_}

    def set_counters(self, **kw):
        c = Counters()
        check(self._L.tmps_get_counters(self._h, C.byref(c)))
        if all(getattr(c, k) == int(v) for k, v in kw.items()):
            return  # already so: no device round trip
        for k, v in kw.items():
            setattr(c, k, int(v))
        check(self._L.tmps_set_counters(self._h, C.byref(c)))

    def last_pressure(self):
        out = np.zeros(6)
        check(self._L.tmps_last_pressure(self._h, _dp(out)))
        return out

    def set_profiling(self, on: bool):
        check(self._L.tmps_set_profiling(self._h, 1 if on else 0))

    def set_launch_mode(self, direct: bool):
        """Graph replay (default) or direct kernel launches for tmps_run."""
        check(self._L.tmps_set_launch_mode(self._h, 1 if direct else 0))

    def adjoint_forward(self, nsteps, every, dt):
        """Setup force plus `nsteps` FP64 steps with a checkpoint every `every`
        steps; returns (final pressure tensor [6], coordination number)."""
        t = np.zeros(6)
        z = C.c_double(0.0)
        check(self._L.tmps_adjoint_forward(self._h, int(nsteps), int(every), float(dt), _dp(t),
                                           C.byref(z)))
        return t, z.value

    def adjoint_backward(self, dL_dtensor, n, initial_state=False):
        """dL/d(kn, kt, gamman, gammat, xmu), dL/d(radius) and, when asked,
        dL/d(initial x, v, omega), for dL/d(final pressure tensor)."""
        g = np.ascontiguousarray(np.asarray(dL_dtensor, dtype=np.float64).reshape(6))
        gc = np.zeros(5)
        gr = np.zeros(n)
        gx = gv = gw = None
        if initial_state:
            gx, gv, gw = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        nullp = C.POINTER(C.c_double)()
        check(self._L.tmps_adjoint_backward(
            self._h, _dp(g), _dp(gc), _dp(gr), _dp(gx) if gx is not None else nullp,
            _dp(gv) if gv is not None else nullp, _dp(gw) if gw is not None else nullp))
        return gc, gr, (gx, gv, gw)

    def kernel_times(self):
        """Average ms and counts per kernel class: kick, step pass, segment
        tail, step close."""
        ms = np.zeros(4)
        cnt = np.zeros(4, dtype=np.int32)
        check(self._L.tmps_kernel_times(self._h, _dp(ms), _ip(cnt)))
        return ms, cnt


class NativeBatch:
    """One tmps_batch: an ensemble of FP32-split systems advanced together.
    The member NativeSystems must outlive it."""

    SEG_LENS = (8, 8, 16, 16, 32, 32)

    def __init__(self, members):
        L = load()
        self._L = L
        self._members = list(members)
        arr = (C.c_void_p * len(self._members))(*[m._h.value for m in self._members])
        h = C.c_void_p()
        check(L.tmps_batch_create(arr, len(self._members), C.byref(h)))
        self._h = h

    def run(self, nsteps, dt, run_pre, autotune=False):
        check(self._L.tmps_batch_run(self._h, int(nsteps), float(dt), 1 if run_pre else 0,
                                     1 if autotune else 0))

    def tuning(self):
        """(segment length, block order, atom-steps/s per configuration,
        member steps redone after rebuild requests in the last run, redo
        passes of the last run, their total wall ms)."""
        seg, order, reb, passes = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
        rates, redo_ms = np.zeros(6), C.c_double(0.0)
        check(self._L.tmps_batch_tuning(self._h, C.byref(seg), C.byref(order), _dp(rates),
                                        C.byref(reb), C.byref(passes), C.byref(redo_ms)))
        return seg.value, order.value, rates, reb.value, passes.value, redo_ms.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.tmps_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NativeWall:
    """A fix wall/gran's device history (tmps_wall_*), bound to one system."""

    def __init__(self, system: "NativeSystem", p: WallParams):
        self._L = load()
        self._sys = system  # keeps the system alive while the wall exists
        self._h = C.c_void_p()
        check(self._L.tmps_wall_create(system.handle, C.byref(p), C.byref(self._h)))

    def reset(self):
        check(self._L.tmps_wall_reset(self._h))

    def post_force(self, dt, coord, vwall_n, vwall_s, history_update):
        check(self._L.tmps_wall_post_force(self._h, float(dt), float(coord), float(vwall_n),
                                           float(vwall_s), 1 if history_update else 0))

    def history(self, n):
        sh = np.zeros((max(n, 1), 3))
        tc = np.zeros(max(n, 1), dtype=np.int32)
        check(self._L.tmps_wall_get_history(self._h, _dp(sh), _ip(tc), int(n)))
        return sh[:n], tc[:n]

    def close(self):
        if self._h:
            self._L.tmps_wall_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(load().tmps_device_name(buf, 256))
    return buf.value.decode()
