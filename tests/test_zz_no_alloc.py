"""tmps_fabric_noalloc, the entry point that allocates nothing (its scratch
is a fixed device array), against tmps_fabric through the C ABI: the same
phi and contact count bit for bit, on two calls in a row (the second on a
smaller system, so the scratch holds the larger call's partials), for both
precisions. GPU."""

import ctypes as C

import numpy as np
import pytest

import simhelpers as H
from taichimps_amd import _native, workloads

pytestmark = pytest.mark.gpu


def _both(sim):
    L = _native.load()
    h = sim.atom._native._h
    dp = C.POINTER(C.c_double)
    phi_a, phi_b = np.zeros(6), np.zeros(6)
    na, nb = C.c_int(-1), C.c_int(-1)
    assert L.tmps_fabric(h, phi_a.ctypes.data_as(dp), C.byref(na)) == 0
    assert L.tmps_fabric_noalloc(h, phi_b.ctypes.data_as(dp), C.byref(nb)) == 0
    return phi_a, na.value, phi_b, nb.value


@pytest.mark.parametrize("precision", ["f64", "f32split"])
def test_fabric_noalloc_equals_fabric_on_two_calls(precision):
    sims = []
    for side in (36, 11):  # 46,656 atoms, then 1,331
        w = workloads.lattice(side, side, side, 1.0, 70 + side, "uniform")
        sim = H.device_sim(w, style="hertz/history", precision=precision, skin=0.05)
        sim.run(20)
        sims.append(sim)
    for sim in sims:
        phi_a, na, phi_b, nb = _both(sim)
        assert na == nb > 0
        assert np.array_equal(phi_a, phi_b), (phi_a, phi_b)
        assert abs(phi_a[:3].sum() - 1.0) < 1e-12  # unit trace
