"""Per-contact computes: ComputeContactAtom, ComputeFabric, ComputeStressAtom
(reference: computes/contact_atom.py, fabric.py, stress_atom.py).

CPU tests pin the oracle's restatement by things other than its own formula:
an O(N^2) enumeration of overlapping pairs, the closed form of a touching
simple-cubic lattice (6 contacts per atom, fabric diag(1/3)), the unit trace
of the fabric tensor, the reference's serial summation order, and the
identity sum(stress) = -(virial tensor) with the virial tree pinned elsewhere.
GPU tests compare the HIP kernels with the oracle: bit for bit in FP64, and
with the stated tolerances for FP32-split systems.
"""

from __future__ import annotations

import numpy as np
import pytest

import simhelpers as H
from oracle import oracle as O
from taichimps_amd import workloads


def _dense(n=6, a=1.0, seed=11):
    return workloads.lattice(n, n, n, a, seed, "uniform")


def _oracle_after_run(w, style="hooke/history", steps=40):
    o = H.oracle_sim(w, style=style)
    assert o.run(steps, 1e-4) == 0
    return o


def _brute_contacts(x, r, lo, hi):
    """Overlapping pairs (i < j, ascending) by enumeration over all pairs,
    minimum image in the periodic box: (i, j, d)."""
    prd = np.asarray(hi) - np.asarray(lo)
    out = []
    n = len(x)
    for i in range(n):
        d = x[i] - x[i + 1:]
        d = d - prd * np.round(d / prd)
        rsq = np.einsum("ij,ij->i", d, d)
        radsum = r[i] + r[i + 1:]
        hit = np.nonzero((rsq < radsum * radsum) & (rsq > 1e-18))[0]
        for k in hit:
            out.append((i, i + 1 + k, d[k]))
    return out


# ------------------------------------------------------------------ CPU pins

def test_contact_atom_is_zero_before_the_first_build():
    o = H.oracle_sim(_dense(3))
    assert np.array_equal(o.contact_atom(), np.zeros(27, dtype=np.int32))
    phi, nc = o.fabric()
    assert nc == 0 and np.array_equal(phi, np.zeros(6))


def test_contact_atom_equals_brute_force_overlap_count():
    """The two-sided tally of the half list equals an all-pairs enumeration
    at the drifted positions of a run (a one-sided tally gives half)."""
    w = _dense()
    o = _oracle_after_run(w)
    st = o.atoms()
    lo, hi, _ = o.box()
    expect = np.zeros(len(st["x"]), dtype=np.int64)
    pairs = _brute_contacts(st["x"], w["radius"], lo, hi)
    assert len(pairs) > 100
    for i, j, _ in pairs:
        expect[i] += 1
        expect[j] += 1
    assert np.array_equal(o.contact_atom(), expect)


def test_touching_cubic_lattice_closed_form():
    """Simple-cubic lattice, spacing 1, radius 0.52: each atom touches its 6
    face neighbours only; Nc = 3N and the fabric is diag(1/3)."""
    n = 5
    g = (np.arange(n) + 0.5)
    x = np.array([[a, b, c] for a in g for b in g for c in g])
    N = len(x)
    o = O.OracleSim(N)
    r = np.full(N, 0.52)
    o.set_atoms(x, np.zeros((N, 3)), np.zeros((N, 3)), r, O.mass_from_radius(r, 1.0))
    o.set_box((0.0,) * 3, (float(n),) * 3, (1, 1, 1))
    o.set_neigh(0.1)
    assert o.nlist_build() == 0
    assert np.array_equal(o.contact_atom(), np.full(N, 6))
    phi, nc = o.fabric()
    assert nc == 3 * N
    np.testing.assert_allclose(phi[:3], 1.0 / 3.0, rtol=0, atol=1e-15)
    assert np.array_equal(phi[3:], np.zeros(3))


def test_fabric_matches_the_serial_sum_and_has_unit_trace():
    """The oracle's tree-ordered sums agree with the reference's serial
    accumulation (fabric.py:43-66) to rounding; each contact adds a unit
    vector's outer product, so trace(phi) = 1."""
    w = _dense(seed=5)
    o = _oracle_after_run(w, style="hertz/history")
    st = o.atoms()
    lo, hi, _ = o.box()
    s = np.zeros(6)
    pairs = _brute_contacts(st["x"], w["radius"], lo, hi)
    for _, _, d in pairs:
        nv = d / np.sqrt(d @ d)
        s += [nv[0] * nv[0], nv[1] * nv[1], nv[2] * nv[2], nv[0] * nv[1], nv[0] * nv[2],
              nv[1] * nv[2]]
    phi, nc = o.fabric()
    assert nc == len(pairs)
    np.testing.assert_allclose(phi, s * (1.0 / nc), rtol=0, atol=1e-13)
    assert abs(phi[:3].sum() - 1.0) < 1e-13


@pytest.mark.parametrize("keflag", [False, True])
def test_stress_atom_is_the_negated_virial_and_sums_to_minus_the_tensor(keflag):
    w = _dense(seed=9)
    o = _oracle_after_run(w, style="hooke/history", steps=25)
    st = o.atoms()
    m = O.mass_from_radius(w["radius"], w["density"])
    v = st["v"]
    expect = -st["virial"]
    if keflag:
        kin = np.stack([m * v[:, 0] * v[:, 0], m * v[:, 1] * v[:, 1], m * v[:, 2] * v[:, 2],
                        m * v[:, 0] * v[:, 1], m * v[:, 0] * v[:, 2], m * v[:, 1] * v[:, 2]],
                       axis=1)
        expect = expect - kin
    stress, total = o.stress_atom(keflag)
    assert np.array_equal(stress, expect)
    # negation is exact, so the same pairwise tree gives -(virial tensor) bitwise
    assert np.array_equal(total, -o.virial(keflag))


# ------------------------------------------------------------------ GPU parity

def _device_computes(sim):
    import taichimps_amd as tm
    atom, dom, nl = sim.atom, sim.domain, sim.neighbor
    cnt = tm.ComputeContactAtom(atom.max_atoms).compute(atom, dom, nl)
    phi, nc = tm.ComputeFabric().compute(atom, dom, nl)
    sa = tm.ComputeStressAtom(atom.max_atoms)
    pair_stress = sa.compute(atom, dom, kinetic=False)
    full_stress = sa.compute(atom, dom, kinetic=True)
    macro = sa.compute_macro(atom, dom, kinetic=True)
    return cnt, phi, nc, pair_stress, full_stress, macro


@pytest.mark.gpu
@pytest.mark.parametrize("style", ["hooke/history", "hertz/history", "granular"])
def test_computes_bitexact_fp64_after_a_run(style):
    w = _dense(8, 1.0, seed=21)
    sim = H.device_sim(w, style=style)
    sim.run(60)
    o = H.oracle_sim(w, style=style)
    assert o.run(60, 1e-4) == 0
    cnt, phi, nc, pair_stress, full_stress, macro = _device_computes(sim)
    assert np.array_equal(cnt, o.contact_atom())
    ophi, onc = o.fabric()
    assert nc == onc and nc > 0
    assert np.array_equal(phi, ophi)
    s0, _ = o.stress_atom(False)
    s1, t1 = o.stress_atom(True)
    assert np.array_equal(pair_stress, s0)
    assert np.array_equal(full_stress, s1)
    lo, hi, _ = o.box()
    assert np.array_equal(macro, t1 / float(np.prod(hi - lo)))
    # the computes only read: the run continues bitwise with the oracle's
    sim.run(20)
    assert o.run(20, 1e-4) == 0
    assert np.array_equal(H.device_state(sim)["x"], o.atoms()["x"])


@pytest.mark.gpu
def test_computes_bitexact_fp64_after_stage_steps():
    """Through Simulation.step (the stage kernels) instead of the fused run."""
    w = _dense(6, 1.0, seed=4)
    sim = H.device_sim(w, style="hertz/history")
    sim.setup_run(15)
    sim.init_simulation()
    for _ in range(15):
        sim.step()
    o = H.oracle_sim(w, style="hertz/history")
    assert o.run(15, 1e-4) == 0
    cnt, phi, nc, pair_stress, _, _ = _device_computes(sim)
    assert np.array_equal(cnt, o.contact_atom())
    assert np.array_equal(phi, o.fabric()[0]) and nc == o.fabric()[1]
    assert np.array_equal(pair_stress, o.stress_atom(False)[0])


@pytest.mark.gpu
def test_computes_fp32_split_within_tolerance():
    """FP32-split system vs the FP64 oracle after 100 steps. Stated
    tolerances: contact counts equal except for pairs whose overlap is within
    FP32 rounding of zero (at most 0.5 % of the atoms differ, total within
    0.5 %); fabric within 1e-4 absolute; macro stress within 1e-3 of its
    largest component."""
    w = _dense(10, 1.0, seed=8)
    sim = H.device_sim(w, style="hertz/history", precision="f32split")
    sim.run(100)
    o = H.oracle_sim(w, style="hertz/history")
    assert o.run(100, 1e-4) == 0
    cnt, phi, nc, _, _, macro = _device_computes(sim)
    ocnt = o.contact_atom()
    n = len(ocnt)
    assert np.count_nonzero(cnt != ocnt) <= max(1, n // 200)
    assert abs(int(cnt.sum()) - int(ocnt.sum())) <= max(2, ocnt.sum() // 200)
    ophi, onc = o.fabric()
    assert np.abs(phi - ophi).max() < 1e-4
    _, t1 = o.stress_atom(True)
    lo, hi, _ = o.box()
    omacro = t1 / float(np.prod(hi - lo))
    assert H.max_rel(macro, omacro) < 1e-3
