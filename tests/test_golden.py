"""Golden fixtures (tests/golden/, written by make_golden.py from the FP64
oracle): the oracle must keep reproducing them (CPU), and the HIP path must
reproduce them bit for bit in FP64 (GPU)."""

from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

import simhelpers as H
from golden import make_golden as G
from oracle import oracle as O
from taichimps_amd import workloads

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _json(name):
    with open(os.path.join(GOLD, name)) as fh:
        return json.load(fh)


def _npz(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


# ------------------------------------------------------------------ oracle

_FLOOR_SEEDS = tuple(range(1, 9))


def g3_fp64_noise_floor():
    """Relative mean-pressure difference, per 100-step chunk of G3, between
    the FP64 oracle run and FP64 runs whose input positions carry U(+-1e-7)
    noise (about the rounding of a position kept as an FP32 offset within a
    bin): the worst over eight noise seeds, made monotone over the chunks.
    This is how far two FP64 runs that start one rounding apart drift in this
    dense Hooke packing."""
    gold = _json("g3_isotropic_8k.json")
    ref = np.array([np.mean(r["pressure"][:3]) for r in gold["series"]])
    w = workloads.case("isotropic_hooke_8k")
    worst = np.zeros(len(ref))
    for seed in _FLOOR_SEEDS:
        wp = dict(w)
        rng = np.random.default_rng(seed)
        wp["x"] = w["x"] + rng.uniform(-1e-7, 1e-7, size=w["x"].shape)
        o = H.oracle_sim(wp, style="hooke/history")
        for k in range(len(ref)):
            assert o.run(G.ISO8K_RUN["chunk_steps"], 1e-4, run_pre=k == 0) == 0
            p = o.last_pressure_tensor()[:3].mean()
            worst[k] = max(worst[k], abs(p - ref[k]) / abs(ref[k]))
    return np.maximum.accumulate(worst)


def test_g3_fp64_noise_floor_shows_the_packing_is_chaotic():
    """Inputs one FP32 rounding apart: the FP64 runs agree to 1e-5 after 100
    steps and disagree by more than 1e-4 after 1,000, the floor the FP32-split
    G3 check is held to."""
    floor = g3_fp64_noise_floor()
    assert floor[0] < 1e-5
    assert floor[-1] > 1e-4


def test_oracle_reproduces_g1_two_sphere():
    gold = _json("g1_two_sphere.json")
    got = G.make_g1()
    assert set(got) == set(gold)
    for name, ref in gold.items():
        for k in ("f", "torque", "virial", "shear"):
            assert np.array_equal(np.asarray(got[name][k]), np.asarray(ref[k])), (name, k)
        assert got[name]["partner"] == ref["partner"]
    # the cases cover stick and slip history and a spin-driven contact
    assert np.abs(np.asarray(gold["hooke_stick"]["shear"])).max() > 0
    assert np.abs(np.asarray(gold["hooke_spin"]["torque"])).max() > 0


def test_oracle_reproduces_g2_lattice27():
    gold = _npz("g2_lattice27.npz")
    got = G.make_g2()
    assert set(got) == set(gold.files)
    for k in gold.files:
        assert np.array_equal(got[k], gold[k]), k
        assert np.abs(gold[k]).max() > 0, k


def test_oracle_reproduces_g3_isotropic_8k():
    gold = _json("g3_isotropic_8k.json")
    got = G.make_g3()
    assert got == gold


@pytest.mark.parametrize("name,maker", [("g4_rebuild_heavy.npz", G.make_g4),
                                        ("g5_history_remap.npz", G.make_g5)])
def test_oracle_reproduces_g4_g5(name, maker):
    gold = _npz(name)
    got = maker()
    assert set(got) == set(gold.files)
    for k in gold.files:
        assert np.array_equal(np.asarray(got[k]), gold[k]), k


def test_g5_history_follows_partner_tags():
    """Across the rebuild every touching pair keeps its shear bits."""
    g = _npz("g5_history_remap.npz")
    before = {(i, j): (p, tuple(s)) for i, j, p, s in
              zip(g["pair_i_before"], g["pair_j_before"], g["partner_before"], g["shear_before"])}
    n_kept = 0
    for i, j, p, s in zip(g["pair_i_after"], g["pair_j_after"], g["partner_after"],
                          g["shear_after"]):
        if p >= 0:
            bp, bs = before[(i, j)]
            assert bp == p and bs == tuple(s)
            n_kept += 1
    assert n_kept == int((g["partner_before"] >= 0).sum()) > 0


# ------------------------------------------------------------------ device

def _device_two_sphere(name, case):
    import taichimps_amd as tm
    style, x, v, om, r, kw, calls = case
    dom = tm.Domain([0.0] * 3, [10.0] * 3, ("p", "p", "p"))
    atom = tm.AtomSystem(max_atoms=2)
    atom.add_particles(x=np.asarray(x, float), v=np.asarray(v, float), omega=np.asarray(om, float),
                       radius=np.asarray(r, float), density=1.0)
    nl = tm.NeighborList(dom, skin=0.1, max_atoms=2, max_neighbors_per_atom=64)
    hist = tm.ContactHistory(2, 64)
    cls = tm.GranHookeHistory if style == "hooke/history" else tm.GranHertzHistory
    pair = cls(dom, kw["kn"], kw["gamman"], kw["kt"], kw["gammat"], kw["xmu"])
    nl.build(atom)
    for _ in range(calls):
        atom.clear_forces()
        pair.compute(atom, nl, hist, 1e-4, True)
    return atom, hist


@pytest.mark.gpu
def test_device_matches_g1_two_sphere():
    gold = _json("g1_two_sphere.json")
    for name, case in G.two_sphere_cases().items():
        atom, hist = _device_two_sphere(name, case)
        ref = gold[name]
        assert np.array_equal(atom.f.to_numpy()[:2], np.asarray(ref["f"])), name
        assert np.array_equal(atom.torque.to_numpy()[:2], np.asarray(ref["torque"])), name
        assert np.array_equal(atom.virial.to_numpy()[:2], np.asarray(ref["virial"])), name
        assert np.array_equal(hist.shear.to_numpy()[:1], np.asarray(ref["shear"])), name


def _device_pair(tm, dom, kind, kw):
    if kind == "legacy":
        cls = tm.GranHookeHistory if kw["style"] == "hooke/history" else tm.GranHertzHistory
        return cls(dom, kw["kn"], kw["gamman"], kw["kt"], kw["gammat"], kw["xmu"])
    return tm.PairGranular(dom, kw["normal"], kw["normal_coeffs"], tangential=kw["tangential"],
                           tangential_coeffs=kw["tangential_coeffs"], damping=kw["damping"])


@pytest.mark.gpu
@pytest.mark.parametrize("label", list(G.LATTICE27_PAIRS))
def test_device_matches_g2_lattice27(label):
    import taichimps_amd as tm
    gold = _npz("g2_lattice27.npz")
    kind, kw = G.LATTICE27_PAIRS[label]
    c = G.lattice27()
    n = len(c["x"])
    for steps in G.LATTICE27_RUN["steps"]:
        dom = tm.Domain([0.0] * 3, [c["box"]] * 3, ("p", "p", "p"))
        atom = tm.AtomSystem(max_atoms=n)
        atom.add_particles(x=c["x"], v=c["v"], omega=c["omega"], radius=c["radius"],
                           density=c["density"])
        nl = tm.NeighborList(dom, skin=G.LATTICE27_RUN["skin"], max_atoms=n,
                             max_neighbors_per_atom=64)
        nl.check = False
        nl.every = 1
        hist = tm.ContactHistory(n, 64)
        sim = tm.Simulation(dom, atom, nl, hist, _device_pair(tm, dom, kind, kw),
                            dt=G.LATTICE27_RUN["dt"])
        sim.add_fix(tm.FixNVESphere(dom))
        if steps == 0:
            sim.init_simulation()
        else:
            sim.run(steps)
        assert np.array_equal(atom.f.to_numpy()[:n], gold[f"{label}_f_{steps}"]), steps
        assert np.array_equal(atom.torque.to_numpy()[:n], gold[f"{label}_torque_{steps}"]), steps


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f64", "f32split"])
def test_device_matches_g3_isotropic_8k(precision):
    gold = _json("g3_isotropic_8k.json")
    w = workloads.case("isotropic_hooke_8k")
    sim = H.device_sim(w, style="hooke/history", precision=precision)
    floor = None
    for k, ref in enumerate(gold["series"]):
        sim.run_pre = k == 0
        sim.run(G.ISO8K_RUN["chunk_steps"])
        p = sim.atom._native.last_pressure()
        if precision == "f64":
            assert np.array_equal(p, np.asarray(ref["pressure"])), k
            sim.domain.pull()
            assert np.array_equal(np.array(sim.domain.boxlo), np.asarray(ref["lo"])), k
            assert sim.neighbor.build_count == ref["build_count"], k
        else:
            # This dense Hooke packing is chaotic at the 1e-7 level (FP64 runs
            # whose input positions differ by U(+-1e-7) drift apart,
            # g3_fp64_noise_floor). FP32-split rounds at that level at every
            # step, not once, and is held to four times that floor (never
            # looser than 1e-5).
            if floor is None:
                floor = g3_fp64_noise_floor()
            mean, mean_ref = p[:3].mean(), np.mean(ref["pressure"][:3])
            tol = max(1e-5, 4.0 * floor[k])
            err = abs(mean - mean_ref) / abs(mean_ref)
            print(f"chunk {k}: split rel err {err:.3e}, FP64 floor {floor[k]:.3e}")
            assert err <= tol, (k, err, floor[k])
    if precision == "f64":
        n = len(w["x"])
        assert _sha(sim.atom.x.to_numpy()[:n]) == gold["x_sha256"]
        assert _sha(sim.atom.v.to_numpy()[:n]) == gold["v_sha256"]
        assert _sha(sim.atom.omega.to_numpy()[:n]) == gold["omega_sha256"]


@pytest.mark.gpu
def test_device_matches_g4_rebuild_heavy():
    gold = _npz("g4_rebuild_heavy.npz")
    R = G.REBUILD_HEAVY_RUN
    w = G.rebuild_heavy_workload()
    sim = H.device_sim(w, style=R["style"], servo=False, skin=R["skin"], dt=R["dt"])
    sim.run(R["steps"])
    nl = sim.neighbor
    n = len(w["x"])
    assert nl.build_count == int(gold["build_count"][0])
    assert np.array_equal(nl.pair_i.to_numpy(), gold["pair_i"])
    assert np.array_equal(nl.pair_j.to_numpy(), gold["pair_j"])
    assert np.array_equal(nl.grid_dim[None], gold["grid"])
    assert np.array_equal(nl.cell_of.to_numpy()[:n], gold["cell_of"])
    assert np.array_equal(sim.history.shear.to_numpy()[: len(gold["shear"])], gold["shear"])
    assert np.array_equal(sim.atom.x.to_numpy()[:n], gold["x"])


@pytest.mark.gpu
def test_device_matches_g5_history_remap():
    gold = _npz("g5_history_remap.npz")
    Hh = G.HISTORY_REMAP_RUN
    w = workloads.lattice(*Hh["dims"], 1.1, Hh["seed"])
    sim = H.device_sim(w, style=Hh["style"], servo=False, skin=Hh["skin"], dt=Hh["dt"],
                       every=10**6)
    sim.run(Hh["steps"])
    hist = sim.history
    m = len(gold["pair_i_before"])
    assert np.array_equal(sim.neighbor.pair_i.to_numpy(), gold["pair_i_before"])
    assert np.array_equal(hist.shear.to_numpy()[:m], gold["shear_before"])
    sim.reneighbor_if_needed(force=True)
    m = len(gold["pair_i_after"])
    assert np.array_equal(sim.neighbor.pair_i.to_numpy(), gold["pair_i_after"])
    assert np.array_equal(sim.neighbor.pair_j.to_numpy(), gold["pair_j_after"])
    assert np.array_equal(hist.partner.to_numpy()[:m], gold["partner_after"])
    assert np.array_equal(hist.shear.to_numpy()[:m], gold["shear_after"])
