"""LAMMPS binary restart files (reference: restart.py, contact_history.py:156-214).

CPU: the writer reproduces, byte for byte, the fixtures G6 that the
reference's own writer made from the same inputs (tests/golden/
make_restart_golden.py), and the reader recovers those inputs from them.
GPU: a run written and read back is unchanged, every contact appears under
both partners (the second copy negated), and a run resumed from a restart
lands bit for bit where the uninterrupted run does.
"""

from __future__ import annotations

import os
import struct
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import restart_case as R  # noqa: E402

from taichimps_amd import restart as RS  # noqa: E402

GOLDEN = os.path.join(HERE, "golden")


def _golden(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return f.read()


def _write_case(tmp_path, full):
    d = R.arrays()
    path = tmp_path / ("full" if full else "plain")
    if full:
        RS.write_restart(path, R.Atom(d), R.Domain(), timestep=R.TIMESTEP, dt=R.DT,
                         history=R.History(d), pair=R.Pair(), pair_style="granular",
                         nlist=R.NeighborList(d))
    else:
        RS.write_restart(path, R.Atom(d), R.Domain(), timestep=R.TIMESTEP, dt=R.DT)
    return path.read_bytes()


@pytest.mark.parametrize("full", [True, False])
def test_writer_matches_the_reference_bytes(tmp_path, full):
    name = "g6_restart_full.bin" if full else "g6_restart_plain.bin"
    assert _write_case(tmp_path, full) == _golden(name)


def test_reader_recovers_the_inputs_from_the_reference_file(tmp_path):
    path = tmp_path / "g6"
    path.write_bytes(_golden("g6_restart_full.bin"))
    data = RS.read_restart(path)
    d = R.arrays()
    assert data.timestep == R.TIMESTEP and data.dt == R.DT and data.natoms == R.N
    assert data.ntypes == 2 and data.units == "si" and data.atom_style == "sphere"
    assert data.boxlo == list(R.BOXLO) and data.boxhi == list(R.BOXHI)
    assert data.boundary == ("p", "p", "f")
    for key in ("x", "v", "omega", "radius", "rmass", "tag", "atom_type"):
        assert np.array_equal(getattr(data, key), d[key]), key
    assert data.pair_style == "granular"
    model = data.pair_settings["models"][0]
    assert model["normal"] == ("hertz/material", [1.0e7, 0.5, 0.3])
    assert model["tangential"] == ("mindlin", [-1.0, 1.0, 0.5])
    assert model["damping"] == ("coeff_restitution", [])
    assert data.fix_peratom_styles == ["NEIGH_HISTORY"]
    # both-sided history: the owner's value under i, negated under the partner
    stored = [dict(zip(*(lambda t, v: (t.tolist(), list(v)))(*RS.parse_neigh_history(r))))
              for r in data.fix_extra]
    idx = {int(t): i for i, t in enumerate(d["tag"])}
    total = 0
    for nc, (i, j, touching) in enumerate(R.PAIRS):
        if not touching:
            continue
        tj, ti = int(d["tag"][j]), int(d["tag"][i])
        assert np.array_equal(stored[i][tj], d["shear"][nc])
        assert np.array_equal(stored[idx[tj]][ti], -d["shear"][nc])
        total += 1
    assert sum(len(s) for s in stored) == 2 * total


def test_plain_file_has_no_history_and_no_pair(tmp_path):
    path = tmp_path / "g6p"
    path.write_bytes(_golden("g6_restart_plain.bin"))
    data = RS.read_restart(path)
    assert data.pair_style == "" and data.fix_peratom_styles == []
    assert all(len(e) == 0 for e in data.fix_extra)
    assert np.array_equal(data.x, R.arrays()["x"])


def test_reader_errors(tmp_path):
    raw = _golden("g6_restart_full.bin")
    bad = tmp_path / "bad"
    bad.write_bytes(b"not a restart" + raw)
    with pytest.raises(ValueError, match="not a LAMMPS restart"):
        RS.read_restart(bad)
    bad.write_bytes(raw[: len(raw) // 2])
    with pytest.raises(ValueError):
        RS.read_restart(bad)
    # the other byte order
    bad.write_bytes(raw[:16] + struct.pack(">i", 1) + raw[20:])
    with pytest.raises(ValueError, match="byte order"):
        RS.read_restart(bad)
    # a newer format revision
    bad.write_bytes(raw[:20] + struct.pack("<i", 4) + raw[24:])
    with pytest.raises(ValueError, match="revision"):
        RS.read_restart(bad)


def test_writer_rejects_a_pair_it_cannot_store(tmp_path):
    d = R.arrays()
    with pytest.raises(ValueError, match="not implemented"):
        RS.write_restart(tmp_path / "x", R.Atom(d), R.Domain(), timestep=0, dt=1e-4,
                         pair=R.Pair(), pair_style="gran/hooke/history")


def test_parse_neigh_history_layout():
    rec = np.array([10.0, 2.0, *RS._int_to_bits([17]), 1.0, 2.0, 3.0,
                    *RS._int_to_bits([23]), -4.0, 5.0, -6.0])
    rec[0] = len(rec)
    tags, vals = RS.parse_neigh_history(rec)
    assert tags.tolist() == [17, 23]
    assert np.array_equal(vals, [[1.0, 2.0, 3.0], [-4.0, 5.0, -6.0]])
    t0, v0 = RS.parse_neigh_history(np.zeros(0))
    assert t0.size == 0 and v0.shape == (0, 3)


# ------------------------------------------------------------------ GPU

def _packing(seed=5):
    from taichimps_amd import workloads
    return workloads.lattice(5, 5, 5, 0.98, seed, "uniform")


def _make_sim(w, style, n_max=None, history=None, rmass=None, tag=None):
    import taichimps_amd as tm
    import simhelpers as H
    n = len(w["x"])
    n_max = n_max or n + 8
    dom = tm.Domain(w["boxlo"], w["boxhi"], ("p", "p", "p"))
    atom = tm.AtomSystem(max_atoms=n_max)
    atom.add_particles(x=w["x"], radius=w["radius"], density=w["density"], v=w["v"],
                       omega=w["omega"], tag=w["tag"] if tag is None else tag)
    if rmass is not None:
        atom.rmass.from_numpy(rmass)
    nl = tm.NeighborList(dom, skin=0.1, max_atoms=n_max, max_neighbors_per_atom=64)
    hist = history if history is not None else tm.ContactHistory(n_max, 64)
    sim = tm.Simulation(dom, atom, nl, hist, H.pair_for(tm, dom, style), dt=1e-4)
    sim.add_fix(tm.FixNVESphere(dom))
    return sim


def _sim_from_restart(path, style):
    import taichimps_amd as tm
    data = RS.read_restart(path)
    n_max = data.natoms + 8
    hist = tm.ContactHistory(n_max, 64)
    if any(len(e) for e in data.fix_extra):
        dec = [RS.parse_neigh_history(e) for e in data.fix_extra]
        hist.load_restart([t for t, _ in dec], [v for _, v in dec])
    w = {"x": data.x, "v": data.v, "omega": data.omega, "radius": data.radius,
         "density": data.density, "tag": data.tag, "boxlo": data.boxlo, "boxhi": data.boxhi}
    sim = _make_sim(w, style, n_max=n_max, history=hist, rmass=data.rmass)
    sim.timestep = data.timestep
    return sim, data


def _write(sim, path, style):
    RS.write_restart(path, sim.atom, sim.domain, timestep=sim.timestep, dt=sim.dt,
                     history=sim.history,
                     pair=sim.pair_style if style == "granular" else None,
                     pair_style=style, nlist=sim.neighbor)


@pytest.mark.gpu
def test_device_round_trip_is_bit_exact(tmp_path):
    sim = _make_sim(_packing(), "granular")
    sim.run(60)
    path = tmp_path / "rt.restart"
    _write(sim, path, "granular")
    data = RS.read_restart(path)
    n = sim.atom.nlocal
    assert data.timestep == sim.timestep and data.natoms == n
    for key in ("x", "v", "omega", "radius", "rmass", "tag"):
        assert np.array_equal(getattr(data, key), getattr(sim.atom, key).to_numpy()[:n]), key
    own = sim.history.touching_by_atom(sim.neighbor, n)
    tags = sim.atom.tag.to_numpy()[:n]
    idx = {int(t): i for i, t in enumerate(tags)}
    stored = []
    for rec in data.fix_extra:
        t, v = RS.parse_neigh_history(rec)
        stored.append({int(a): v[k] for k, a in enumerate(t)})
    total = 0
    for i in range(n):
        for ptag, value in own[i]:
            assert np.array_equal(stored[i][ptag], value)
            assert np.array_equal(stored[idx[ptag]][int(tags[i])], -value)
            total += 1
    assert total > 20, "the packing should hold contacts"
    assert sum(len(s) for s in stored) == 2 * total


@pytest.mark.gpu
@pytest.mark.parametrize("style", ["granular", "hooke/history"])
def test_resumed_run_equals_the_uninterrupted_run_bitwise(tmp_path, style):
    """A restart resumes as a second `run` command does: history, positions
    and velocities carried over, the same trajectory bit for bit."""
    w = _packing(seed=9)
    cont = _make_sim(w, style)
    cont.run(80)
    cont.run(70)
    first = _make_sim(w, style)
    first.run(80)
    path = tmp_path / "mid.restart"
    _write(first, path, style)
    second, data = _sim_from_restart(path, style)
    assert sum(len(e) > 2 for e in data.fix_extra) > 10, "the restart should carry history"
    second.run(70)
    n = cont.atom.nlocal
    for key in ("x", "v", "omega", "f", "torque"):
        a = getattr(cont.atom, key).to_numpy()[:n]
        b = getattr(second.atom, key).to_numpy()[:n]
        assert np.array_equal(a, b), f"{key}: max abs {np.abs(a - b).max():.3e}"
    assert second.timestep == cont.timestep


@pytest.mark.gpu
def test_lost_history_changes_the_trajectory(tmp_path):
    """Control: the same resume without load_restart diverges, so the
    bitwise test above depends on the carried history."""
    import taichimps_amd as tm
    w = _packing(seed=9)
    cont = _make_sim(w, "granular")
    cont.run(80)
    cont.run(70)
    first = _make_sim(w, "granular")
    first.run(80)
    path = tmp_path / "mid.restart"
    _write(first, path, "granular")
    data = RS.read_restart(path)
    w2 = {"x": data.x, "v": data.v, "omega": data.omega, "radius": data.radius,
          "density": data.density, "tag": data.tag, "boxlo": data.boxlo, "boxhi": data.boxhi}
    second = _make_sim(w2, "granular", n_max=data.natoms + 8,
                       history=tm.ContactHistory(data.natoms + 8, 64), rmass=data.rmass)
    second.timestep = data.timestep
    second.run(70)
    n = cont.atom.nlocal
    assert not np.array_equal(cont.atom.x.to_numpy()[:n], second.atom.x.to_numpy()[:n])


def test_history_with_no_touching_contacts_round_trips(tmp_path):
    """Every atom's FixNeighHistory record is [2, 0]: written and read back."""
    d = R.arrays()
    for k in ("pair_i", "pair_j", "partner"):
        d[k] = d[k][:0]
    d["shear"] = d["shear"][:0]
    path = tmp_path / "empty"
    RS.write_restart(path, R.Atom(d), R.Domain(), timestep=1, dt=1e-4, history=R.History(d),
                     nlist=R.NeighborList(d))
    data = RS.read_restart(path)
    assert data.fix_peratom_styles == ["NEIGH_HISTORY"]
    for rec in data.fix_extra:
        assert np.array_equal(rec, [2.0, 0.0])
        t, v = RS.parse_neigh_history(rec)
        assert t.size == 0 and v.shape == (0, 3)


@pytest.mark.gpu
def test_fp32_split_resume_tracks_the_uninterrupted_run():
    """FP32-split systems: the restart stores decoded FP64 positions and the
    FP32 history widened to FP64; a resumed run re-encodes them, so it follows
    the uninterrupted run to FP32 rounding (positions within 1e-4 after 70
    more steps) with the same contacts carried over."""
    import os
    import tempfile

    import taichimps_amd as tm
    import simhelpers as H
    w = _packing(seed=9)
    n = len(w["x"])

    def make(history=None, data=None):
        src = w if data is None else {
            "x": data.x, "v": data.v, "omega": data.omega, "radius": data.radius,
            "density": data.density, "tag": data.tag, "boxlo": data.boxlo,
            "boxhi": data.boxhi}
        dom = tm.Domain(src["boxlo"], src["boxhi"], ("p", "p", "p"))
        atom = tm.AtomSystem(max_atoms=n, float_type="f32split")
        atom.add_particles(x=src["x"], radius=src["radius"], density=src["density"],
                           v=src["v"], omega=src["omega"], tag=src["tag"])
        if data is not None:
            atom.rmass.from_numpy(data.rmass)
        nl = tm.NeighborList(dom, skin=0.1, max_atoms=n, max_neighbors_per_atom=64)
        sim = tm.Simulation(dom, atom, nl, history or tm.ContactHistory(n, 64),
                            H.pair_for(tm, dom, "hooke/history"), dt=1e-4)
        sim.add_fix(tm.FixNVESphere(dom))
        return sim

    cont = make()
    cont.run(80)
    cont.run(70)
    first = make()
    first.run(80)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "mid.restart")
        _write(first, path, "hooke/history")
        data = RS.read_restart(path)
    hist = tm.ContactHistory(n, 64)
    dec = [RS.parse_neigh_history(e) for e in data.fix_extra]
    hist.load_restart([t for t, _ in dec], [v for _, v in dec])
    assert sum(len(t) for t, _ in dec) > 20
    second = make(history=hist, data=data)
    second.timestep = data.timestep
    second.run(70)
    dx = np.abs(cont.atom.x.to_numpy()[:n] - second.atom.x.to_numpy()[:n]).max()
    assert dx < 1e-4, dx
    assert cont.timestep == second.timestep
