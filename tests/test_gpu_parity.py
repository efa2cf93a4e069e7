"""Parity of the HIP path (libtmps_hip.so, through the reference-style API)
with the CPU oracle on the same inputs.

FP64: bins, neighbor sets, history carry-over, per-atom forces/torques/virial
and whole trajectories (positions, velocities, box) are compared bit for bit
(np.array_equal). FP32-split: against the FP64 oracle with the tolerances
stated in each test.
"""

import numpy as np
import pytest

import simhelpers as H
from oracle import oracle as O
from taichimps_amd import workloads

pytestmark = pytest.mark.gpu

STYLES = ["hooke/history", "hertz/history", "granular"]


def dense(n=8, a=1.0, seed=7):
    """Jittered lattice tighter than the radii: most neighbours touch."""
    return workloads.lattice(n, n, n, a, seed, "uniform")


def test_device_is_gfx950():
    from taichimps_amd import _native
    name = _native.device_name()
    print(name)
    assert "gfx950" in name


def _assert_state_equal(dev, orc, keys=("x", "v", "omega", "f", "torque", "virial")):
    for k in keys:
        assert np.array_equal(dev[k], orc[k]), (
            f"{k} differs: max rel {H.max_rel(dev[k], orc[k]):.3e}")


def test_bins_and_half_list_bitexact():
    w = dense(12, 1.0, seed=3)
    sim = H.device_sim(w, servo=False)
    sim.neighbor.build(sim.atom)
    o = H.oracle_sim(w, servo=False)
    assert o.nlist_build() == 0
    g, cs, cell_of = o.grid()
    n = len(w["x"])
    assert np.array_equal(sim.neighbor.grid_dim[None], g)
    assert np.array_equal(sim.neighbor.cell_size[None], cs)
    assert np.array_equal(sim.neighbor.cell_of.to_numpy()[:n], cell_of)
    pi, pj, _, _ = o.pairs()
    assert np.array_equal(sim.neighbor.pair_i.to_numpy(), pi)
    assert np.array_equal(sim.neighbor.pair_j.to_numpy(), pj)
    assert sim.neighbor.npairs[None] == len(pi)


def test_long_rows_take_the_fill_path_bitexact():
    """A skin wide enough that rows exceed the list kernel's 32-entry scratch
    (about 45 partners per atom): the list is rebuilt by the stencil-walk fill
    kernel and still equals the oracle's, and a run over it stays bitwise."""
    w = dense(9, 1.0, seed=9)
    sim = H.device_sim(w, servo=False, skin=1.2)
    sim.neighbor.build(sim.atom)
    o = H.oracle_sim(w, servo=False, skin=1.2)
    assert o.nlist_build() == 0
    pi, pj, _, _ = o.pairs()
    counts = np.bincount(np.concatenate([pi, pj]), minlength=len(w["x"]))
    assert counts.max() > 32
    assert np.array_equal(sim.neighbor.pair_i.to_numpy(), pi)
    assert np.array_equal(sim.neighbor.pair_j.to_numpy(), pj)
    sim2 = H.device_sim(w, style="hertz/history", skin=1.2)
    o2 = H.oracle_sim(w, style="hertz/history", skin=1.2)
    sim2.run(20)
    assert o2.run(20, 1e-4) == 0
    _assert_state_equal(H.device_state(sim2), H.oracle_state(o2))


@pytest.mark.parametrize("n", [200, 1200])
def test_capped_thin_periodic_box_collapses_after_scaling(n):
    """A 40 x 40 x 5 periodic box under a 1,000-bin cap scales to a periodic
    2-bin axis; both the reported grid and the build's own grid collapse it
    to one bin, so the list has each pair once and equals the oracle's (200
    atoms: the build's grid is the capped one itself)."""
    rng = np.random.default_rng(23)
    hi = np.array([40.0, 40.0, 5.0])
    x = rng.uniform(0.0, 1.0, size=(n, 3)) * hi
    r = rng.uniform(0.2, 0.45, size=n)
    r[0] = 0.45
    w = dict(x=x, radius=r, density=np.ones(n), v=np.zeros((n, 3)), omega=np.zeros((n, 3)),
             tag=np.arange(1, n + 1), boxlo=np.zeros(3), boxhi=hi)
    sim = H.device_sim(w, servo=False, max_nb=128)
    sim.neighbor.max_grid_cells = 1000
    sim.neighbor.build(sim.atom)
    o = H.oracle_sim(w, servo=False, max_nb=128)
    o.set_neigh(0.1, 0, 1, True, max_grid_cells=1000)
    assert o.nlist_build() == 0
    g, _, cell_of = o.grid()
    assert g.tolist() == [20, 20, 1]
    assert np.array_equal(sim.neighbor.grid_dim[None], g)
    assert np.array_equal(sim.neighbor.cell_of.to_numpy()[:n], cell_of)
    pi, pj, _, _ = o.pairs()
    dpi, dpj = sim.neighbor.pair_i.to_numpy(), sim.neighbor.pair_j.to_numpy()
    assert len(set(zip(dpi.tolist(), dpj.tolist()))) == len(dpi)
    assert np.array_equal(dpi, pi)
    assert np.array_equal(dpj, pj)


def test_capped_grid_bitexact():
    # a box whose uncapped grid exceeds max_grid_cells, as the 1M-atom case does
    w = dense(10, 1.1, seed=5)
    sim = H.device_sim(w, servo=False)
    sim.neighbor.max_grid_cells = 300
    sim.neighbor.build(sim.atom)
    o = H.oracle_sim(w, servo=False)
    o.set_neigh(0.1, 0, 1, True, max_grid_cells=300)
    assert o.nlist_build() == 0
    g, _, cell_of = o.grid()
    assert int(np.prod(g)) <= 300
    assert np.array_equal(sim.neighbor.grid_dim[None], g)
    assert np.array_equal(sim.neighbor.cell_of.to_numpy()[: len(cell_of)], cell_of)
    pi, pj, _, _ = o.pairs()
    assert np.array_equal(sim.neighbor.pair_i.to_numpy(), pi)
    assert np.array_equal(sim.neighbor.pair_j.to_numpy(), pj)


@pytest.mark.parametrize("style", STYLES)
def test_pair_compute_bitexact(style):
    w = dense(8, 1.0, seed=11)
    w["omega"] = np.random.default_rng(12).normal(0.0, 2.0, size=w["v"].shape)
    w["v"] = w["v"] * 10.0
    sim = H.device_sim(w, style=style, servo=False)
    o = H.oracle_sim(w, style=style, servo=False)
    sim.neighbor.build(sim.atom)
    assert o.nlist_build() == 0
    for shearupdate in (False, True, True):
        sim.atom.clear_forces()
        sim.pair_style.compute(sim.atom, sim.neighbor, sim.history, 1e-4, shearupdate)
        o.clear_forces()
        o.pair_compute(1e-4, shearupdate)
        _assert_state_equal(H.device_state(sim), H.oracle_state(o), ("f", "torque", "virial"))
        _, _, pa, sh = o.pairs()
        assert np.array_equal(sim.history.partner.to_numpy()[: len(pa)], pa)
        assert np.array_equal(sim.history.shear.to_numpy()[: len(pa)], sh)
    f = sim.atom.f.to_numpy()
    assert np.abs(f).max() > 0.0


@pytest.mark.parametrize("style", STYLES)
def test_device_run_bitexact(style):
    """Simulation.run on the device vs the oracle's Simulation.run, 60 steps with the servo."""
    w = dense(9, 1.02, seed=21)
    sim = H.device_sim(w, style=style)
    o = H.oracle_sim(w, style=style)
    sim.run(60)
    assert sim.last_run_device_resident
    assert o.run(60, 1e-4) == 0
    dev, orc = H.device_state(sim), H.oracle_state(o)
    _assert_state_equal(dev, orc)
    assert np.array_equal(dev["boxlo"], orc["boxlo"]) and np.array_equal(dev["boxhi"], orc["boxhi"])
    c = o.counters()
    assert sim.timestep == c["timestep"] == 60
    assert sim.neighbor.build_count == c["build_count"]


@pytest.mark.parametrize("precision", ["f64", "f32split"])
def test_device_run_is_deterministic(precision):
    """Two identical device runs agree bit for bit (no race in the fused passes)."""
    w = dense(10, 1.0, seed=25)
    out = []
    for _ in range(2):
        sim = H.device_sim(w, style="hertz/history", precision=precision)
        sim.run(70)
        out.append(H.device_state(sim))
    for k in ("x", "v", "omega", "f", "torque", "virial", "boxhi"):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_stage_path_bitexact():
    """The reference's stage order driven from Python, one kernel per stage."""
    w = dense(8, 1.02, seed=23)
    sim = H.device_sim(w, style="hertz/history")
    o = H.oracle_sim(w, style="hertz/history")
    sim.setup_run(25)
    sim.init_simulation()
    for _ in range(25):
        sim.step()
    assert o.run(25, 1e-4) == 0
    dev, orc = H.device_state(sim), H.oracle_state(o)
    _assert_state_equal(dev, orc)
    assert np.array_equal(dev["boxhi"], orc["boxhi"])


def test_rebuild_heavy_run_bitexact():
    """Fast atoms: several skin-triggered rebuilds inside one device run."""
    w = dense(7, 1.05, seed=31)
    w["v"] = np.random.default_rng(32).normal(0.0, 4.0, size=w["v"].shape)
    kw = dict(kn=2.0e4, kt=2.0 / 7.0 * 2.0e4, gamman=5.0, gammat=2.5)
    sim = H.device_sim(w, style="hooke/history", dt=5e-4, **kw)
    o = H.oracle_sim(w, style="hooke/history", **kw)
    sim.run(150)
    assert o.run(150, 5e-4) == 0
    c = o.counters()
    assert c["build_count"] >= 4, "the case should rebuild several times"
    assert sim.neighbor.build_count == c["build_count"]
    assert sim.neighbor.last_build_step == c["last_build_step"]
    _assert_state_equal(H.device_state(sim), H.oracle_state(o))


def test_rebuild_every_step_bitexact():
    """neigh_modify every 1 check no: history carried across every step's rebuild."""
    w = dense(6, 1.0, seed=41)
    w["omega"] = np.random.default_rng(42).normal(0.0, 3.0, size=w["v"].shape)
    sim = H.device_sim(w, style="hooke/history", check=False)
    o = H.oracle_sim(w, style="hooke/history", check=False)
    sim.run(12)
    assert o.run(12, 1e-4) == 0
    assert sim.neighbor.build_count == o.counters()["build_count"] == 13
    _assert_state_equal(H.device_state(sim), H.oracle_state(o))


def test_cundall_triaxial_style_bitexact():
    """pair granular + damping/cundall + x,y pressure servo with z erate."""
    w = dense(8, 1.0, seed=51)
    servo_kw = dict(axes={"x": {"style": "pressure", "ptarget": 1.0, "pgain": 1e-4},
                          "y": {"style": "pressure", "ptarget": 1.0, "pgain": 1e-4},
                          "z": {"style": "erate", "rate": -1e-3}})
    sim = H.device_sim(w, style="granular", cundall=(0.1, 0.1), servo_kw=servo_kw)
    o = H.oracle_sim(w, style="granular", cundall=(0.1, 0.1), servo_kw=servo_kw)
    sim.run(40)
    assert o.run(40, 1e-4) == 0
    dev, orc = H.device_state(sim), H.oracle_state(o)
    _assert_state_equal(dev, orc)
    assert np.array_equal(dev["boxhi"], orc["boxhi"])


def test_split_run_vs_fp64_oracle():
    """FP32-split against the FP64 oracle: pressure within 1e-5 relative after 100 steps."""
    w = dense(10, 1.02, seed=61)
    sim = H.device_sim(w, style="hertz/history", precision="f32split")
    o = H.oracle_sim(w, style="hertz/history")
    sim.run(100)
    assert o.run(100, 1e-4) == 0
    p_dev = sim.atom._native.last_pressure()
    p_orc = o.last_pressure_tensor()
    mean_dev, mean_orc = p_dev[:3].mean(), p_orc[:3].mean()
    print("mean P", mean_dev, mean_orc)
    assert abs(mean_dev - mean_orc) <= 1e-5 * abs(mean_orc)
    dev, orc = H.device_state(sim), H.oracle_state(o)
    # positions: FP32 offsets in a dense, fast-moving packing over 100 steps
    assert np.abs(dev["x"] - orc["x"]).max() < 2e-5
    assert H.max_rel(dev["f"], orc["f"]) < 1e-2


# ---------------------------------------------------------------------------
# The reference's own runnable tests, on the device
# ---------------------------------------------------------------------------

def test_reference_hooke_normal_literal():
    """tests/test_hooke.py:16-55 -- LAMMPS-cross-checked literal, rtol 1e-12."""
    import taichimps_amd as tm
    dom = tm.Domain([-10.0] * 3, [10.0] * 3, ("f", "f", "f"))
    atom = tm.AtomSystem(max_atoms=10)
    atom.add_particles(tag=[1, 2], atom_type=[1, 1], x=[[0.0, 0.0, 0.0], [1.9, 0.0, 0.0]],
                       v=[[0.1, 0.0, 0.0], [-0.1, 0.0, 0.0]], radius=[1.0, 1.0],
                       density=[1000.0, 1000.0])
    nl = tm.NeighborList(domain=dom, skin=0.2, max_atoms=10, max_neighbors=10)
    nl.build(atom)
    hist = tm.ContactHistory(max_atoms=10, max_neighbors=10)
    pair = tm.GranHookeHistory(dom, kn=1000.0, gamman=10.0, kt=0.0, gammat=0.0, xmu=0.0, dampflag=0)
    pair.compute(atom, nl, hist, dt=0.001)
    f = atom.f.to_numpy()
    expected = 4288.790204786391
    np.testing.assert_allclose(f[0], [-expected, 0.0, 0.0], rtol=1e-12)
    np.testing.assert_allclose(f[1], [expected, 0.0, 0.0], rtol=1e-12)


def test_reference_skin_displacement_check():
    """tests/test_domain_neighbor.py:58-92."""
    import taichimps_amd as tm
    dom = tm.Domain([-10.0] * 3, [10.0] * 3, ("f", "f", "f"))
    atom = tm.AtomSystem(max_atoms=10)
    atom.add_particles(tag=[1, 2, 3], atom_type=[1, 1, 1],
                       x=[[0.0, 0.0, 0.0], [1.5, 0.0, 0.0], [5.0, 0.0, 0.0]],
                       radius=[1.0, 1.0, 1.0], density=[1000.0] * 3)
    nl = tm.NeighborList(domain=dom, skin=0.2, max_atoms=10, max_neighbors=10)
    assert nl.check_and_build(atom, timestep=0) is True and nl.build_count == 1
    assert nl.num_neighbors.to_numpy()[0] == 1 and nl.neighbors.to_numpy()[0, 0] == 1
    assert nl.check_and_build(atom, timestep=1) is False and nl.build_count == 1
    atom.x[0] = [0.08, 0.0, 0.0]
    assert nl.check_and_build(atom, timestep=2) is False and nl.build_count == 1
    atom.x[0] = [0.12, 0.0, 0.0]
    assert nl.check_and_build(atom, timestep=3) is True and nl.build_count == 2


def test_reference_pbc_wrap_and_friction_sign():
    """tests/test_domain_neighbor.py:15-31 and tests/test_lammps_parity.py:159-190."""
    import taichimps_amd as tm
    dom = tm.Domain([0.0] * 3, [10.0] * 3, ("p", "p", "p"))
    atom = tm.AtomSystem(max_atoms=10)
    atom.add_particles(tag=[1], atom_type=[1], x=[[10.5, -0.2, 5.0]], radius=[0.5],
                       density=[1000.0])
    dom.pbc(atom)
    x = atom.x.to_numpy()[0]
    assert np.isclose(x[0], 0.5) and np.isclose(x[1], 9.8)

    radius, overlap, spin = 0.5, 0.01, 3.0
    d = 2.0 * radius - overlap
    dom = tm.Domain([0.0] * 3, [6.0] * 3, ("p", "p", "p"))
    atom = tm.AtomSystem(max_atoms=8)
    atom.add_particles(x=np.array([[3.0, 3.0, 3.0], [3.0 + d, 3.0, 3.0]]), radius=radius,
                       density=2650.0, omega=np.array([[0.0, 0.0, spin], [0.0, 0.0, 0.0]]))
    nl = tm.NeighborList(domain=dom, max_atoms=8, max_neighbors_per_atom=32, skin=0.1)
    hist = tm.ContactHistory(max_atoms=8, max_neighbors=32)
    pair = tm.GranHookeHistory(dom, kn=1.0e5, gamman=10.0, kt=5.0e4, gammat=5.0, xmu=0.5)
    sim = tm.Simulation(dom, atom, nl, hist, pair, dt=1e-4)
    sim.add_fix(tm.FixNVESphere(dom))
    sim.init_simulation()
    sim.step()
    f = atom.f.to_numpy()[:2]
    t = atom.torque.to_numpy()[:2]
    assert f[0][1] < 0.0 and t[0][2] < 0.0
    np.testing.assert_allclose(f[0] + f[1], np.zeros(3), atol=1e-12 * np.abs(f).max())


def _random_packing(rng, n=64, box=6.0):
    side = int(np.ceil(n ** (1.0 / 3.0)))
    spacing = box / side
    pts = [[(ix + 0.5) * spacing, (iy + 0.5) * spacing, (iz + 0.5) * spacing]
           for ix in range(side) for iy in range(side) for iz in range(side)][:n]
    pts = np.asarray(pts, dtype=np.float64)
    pts += rng.uniform(-0.02 * spacing, 0.02 * spacing, size=pts.shape)
    return pts % box


def _parity_sim(tm, positions, radius, box=6.0, velocities=None, omegas=None):
    dom = tm.Domain([0.0] * 3, [box] * 3, ("p", "p", "p"))
    n = len(positions)
    atom = tm.AtomSystem(max_atoms=max(n, 8))
    atom.add_particles(x=positions, radius=radius, density=2650.0, v=velocities, omega=omegas)
    nl = tm.NeighborList(domain=dom, max_atoms=max(n, 8), max_neighbors_per_atom=32, skin=0.1)
    hist = tm.ContactHistory(max_atoms=max(n, 8), max_neighbors=32)
    pair = tm.GranHookeHistory(dom, kn=1.0e5, gamman=10.0, kt=5.0e4, gammat=5.0, xmu=0.5)
    sim = tm.Simulation(dom, atom, nl, hist, pair, dt=1e-4)
    sim.add_fix(tm.FixNVESphere(dom))
    return sim


def test_reference_pressure_translation_invariance():
    """tests/test_lammps_parity.py:97-121, rtol 1e-9."""
    import taichimps_amd as tm
    rng = np.random.default_rng(7)
    box = 6.0
    radius = 0.52 * (box / 4)
    pts = _random_packing(rng, 64, box)

    def pressure_for(offset):
        sim = _parity_sim(tm, (pts + offset) % box, radius, box)
        sim.init_simulation()
        return tm.Computes().compute_pressure_tensor(sim.atom, sim.domain, kinetic=False)

    p_ref = pressure_for(np.zeros(3))
    p_shift = pressure_for(np.array([0.37 * box, 0.11 * box, 0.53 * box]))
    assert np.linalg.norm(p_ref) > 0.0
    np.testing.assert_allclose(p_shift, p_ref, rtol=1e-9, atol=1e-6 * np.linalg.norm(p_ref))


def test_reference_shear_history_survives_rebuild():
    """tests/test_lammps_parity.py:124-156, rtol 1e-8."""
    import taichimps_amd as tm
    rng = np.random.default_rng(11)
    box = 6.0
    radius = 0.52 * (box / 4)
    pts = _random_packing(rng, 64, box)
    vel = rng.normal(scale=0.05, size=pts.shape)
    omg = rng.normal(scale=0.5, size=pts.shape)

    def run(rebuild_every_step):
        sim = _parity_sim(tm, pts, radius, box, velocities=vel, omegas=omg)
        if rebuild_every_step:
            sim.neighbor.check = False
            sim.neighbor.every = 1
        else:
            sim.neighbor.check = True
            sim.neighbor.every = 10**6
        sim.run(40)
        return sim.atom.f.to_numpy()[: sim.atom.nlocal].copy()

    f_rebuilt = run(True)
    f_kept = run(False)
    scale = np.abs(f_kept).max()
    assert scale > 0.0
    np.testing.assert_allclose(f_rebuilt, f_kept, rtol=1e-8, atol=1e-8 * scale)


def test_reference_remap_x_keeps_fractional_positions():
    """tests/test_deform_pressure.py:223-255 (atol 1e-12) and test_deform_remap_x.py (rel 1e-9)."""
    import taichimps_amd as tm
    BOX, DT = 0.02, 1.0e-6
    dom = tm.Domain([0.0] * 3, [BOX] * 3, ("p", "p", "p"))
    atom = tm.AtomSystem(max_atoms=2)
    atom.add_particles(x=np.array([[0.5 * BOX] * 3, [0.25 * BOX] * 3]), radius=1e-4, density=2650.0)
    fix = tm.FixDeformPressure(dom, axes={d: {"style": "erate", "rate": -1.0e3} for d in "xyz"},
                               nevery=1)
    atom._bind_domain(dom)
    fix.setup(0)
    for _ in range(50):
        fix.end_of_step(atom, DT)
    dom.pull()
    prd, lo = np.array(dom.prd), np.array(dom.boxlo)
    pos = atom.x.to_numpy()[:2]
    assert prd[0] < BOX
    np.testing.assert_allclose((pos[0] - lo) / prd, [0.5] * 3, atol=1e-12)
    np.testing.assert_allclose((pos[1] - lo) / prd, [0.25] * 3, atol=1e-12)
    assert prd[2] == pytest.approx(BOX * (1.0 - 1.0e3 * 50 * DT), rel=1e-9)


def test_large_system_reduction_levels_bitexact():
    """More than 4096 blocks of 256 atoms (1,061,208 atoms): the step close
    first reduces aligned groups of 4096 block partials in parallel. The
    servo-driven trajectory and box stay bitwise the oracle's."""
    w = workloads.lattice(102, 102, 102, 1.1, 1002, "uniform")
    assert len(w["x"]) > 4096 * 256
    sim = H.device_sim(w, style="hertz/history")
    o = H.oracle_sim(w, style="hertz/history")
    sim.run(4)
    assert sim.last_run_device_resident
    assert o.run(4, 1e-4) == 0
    dev, orc = H.device_state(sim), H.oracle_state(o)
    _assert_state_equal(dev, orc, ("x", "v", "omega"))
    assert np.array_equal(dev["boxlo"], orc["boxlo"]) and np.array_equal(dev["boxhi"], orc["boxhi"])


def test_chained_runs_match_the_stage_path():
    """Device-resident runs of one step each, chained without a new fix
    setup, follow the stage path bit for bit through list rebuilds that grow
    the list: the segment tail copies every slot's history even when its
    captured graph predates the rebuild."""
    from taichimps_amd import _native
    w = dense(6, 1.0, seed=43)
    rng = np.random.default_rng(44)
    w["v"] = rng.normal(0.0, 3.0, size=w["v"].shape)
    w["omega"] = rng.normal(0.0, 5.0, size=w["v"].shape)
    a = H.device_sim(w, style="hertz/history", skin=0.02)
    b = H.device_sim(w, style="hertz/history", skin=0.02)
    b.setup_run(30)
    b.init_simulation()
    a.setup_run(30)
    N = a.atom._native
    a.neighbor._bind(a.atom)
    a.neighbor._push()
    a.history._bind(a.atom)
    a.pair_style._install(a.atom)
    N.set_counters(timestep=0, setup_done=0)
    for step in range(30):
        N.run(1, a.dt, run_pre=step == 0)
        b.step()
    c = N.counters()
    a.timestep = c["timestep"]
    a.neighbor._pull()
    a.domain.mark_device_authoritative(a.atom)
    assert a.neighbor.build_count >= 3
    sa, sb = H.device_state(a), H.device_state(b)
    _assert_state_equal(sa, sb, ("x", "v", "omega", "f", "torque"))
    np_ = a.neighbor.npairs[None]
    assert np.array_equal(a.history.shear.to_numpy()[:np_], b.history.shear.to_numpy()[:np_])
    assert _native is not None


@pytest.mark.parametrize("precision", ["f64", "f32split"])
def test_run_on_a_caller_stream_with_prepared_graphs_is_bitwise_the_plain_run(precision):
    """tmps_prepare_run captures a run's segment graphs ahead of it and
    tmps_set_stream moves the system onto a caller's stream (here a torch
    stream, which HIP events can bracket): the run's results do not change,
    rebuilds included, and a following run on the system's own stream
    continues it. The getters after the run (counters, box, pressure) read the
    run's end state."""
    torch = pytest.importorskip("torch")
    w = dense(9, 1.0, seed=61)
    rng = np.random.default_rng(62)
    w["v"] = rng.normal(0.0, 1.5, size=w["v"].shape)
    a = H.device_sim(w, style="hertz/history", precision=precision, skin=0.02)
    b = H.device_sim(w, style="hertz/history", precision=precision, skin=0.02)
    for s in (a, b):
        s.run(3)
        s.run_pre = False
    a.run(45)
    a.run(13)
    b.prepare_run(45)
    stream = torch.cuda.Stream()
    b.atom._native.set_stream(stream.cuda_stream)
    assert b.atom._native.stream() == stream.cuda_stream
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    b.run(45)
    stop.record(stream)
    torch.cuda.synchronize()
    assert start.elapsed_time(stop) > 0.0
    b.atom._native.set_stream(None)
    b.run(13)
    assert a.neighbor.build_count == b.neighbor.build_count > 3
    sa, sb = H.device_state(a), H.device_state(b)
    _assert_state_equal(sa, sb, ("x", "v", "omega", "f", "torque"))
    assert np.array_equal(sa["boxhi"], sb["boxhi"])
    assert np.array_equal(a.atom._native.last_pressure(), b.atom._native.last_pressure())
    assert a.atom._native.counters() == b.atom._native.counters()


def test_device_run_whose_list_outgrows_its_buffers_bitexact():
    """A lattice of spacing 1.6 (no pair within the cutoff at the start)
    compressed by erate on every axis to spacing ~1.0: within one
    device-resident run the list grows from empty to about six entries per
    atom, past the buffers sized at the run's first build (the need, a
    quarter more and 1,024 slots). A rebuild decided on the device then finds
    the list too long for them; the run redoes that step on the host, grows
    the buffers and goes on on the device, its rebuild chain captured again
    for the new buffers. A second run continues it. Both equal the oracle bit
    for bit."""
    w = workloads.lattice(10, 10, 10, 1.6, 71, "uniform")
    rng = np.random.default_rng(72)
    w["v"] = rng.normal(0.0, 0.5, size=w["v"].shape)
    spec = dict(axes={d: {"style": "erate", "rate": -15.0} for d in "xyz"})
    sim = H.device_sim(w, style="hertz/history", servo=True, skin=0.05, servo_kw=spec)
    o = H.oracle_sim(w, style="hertz/history", servo=False, skin=0.05)
    o.set_servo(O.servo_params(**spec))
    for steps in (250, 60):
        sim.run(steps)
        assert sim.last_run_device_resident
        assert o.run(steps, 1e-4) == 0
        dev, orc = H.device_state(sim), H.oracle_state(o)
        _assert_state_equal(dev, orc)
        assert np.array_equal(dev["boxlo"], orc["boxlo"])
        assert np.array_equal(dev["boxhi"], orc["boxhi"])
        assert sim.neighbor.build_count == o.counters()["build_count"]
    # the list outgrew the first build's buffers (1,024 slots: nothing touched then)
    assert 2 * sim.neighbor.npairs[None] > 2048


@pytest.mark.parametrize("check", [True, False])
@pytest.mark.parametrize("precision", ["f64", "f32split"])
def test_graph_and_direct_launches_give_the_same_run(precision, check):
    """tmps_run replays captured graphs (segments, and the rebuild chain with
    its step read on the device) or launches every kernel itself: over runs
    with rebuilds the two give the same bits. With `check no` every step is a
    rebuild, so every run also ends on a redone step (which stores the
    virial instead of applying the next half-kick)."""
    w = dense(9, 1.0, seed=81)
    rng = np.random.default_rng(82)
    w["v"] = rng.normal(0.0, 1.5, size=w["v"].shape)
    a = H.device_sim(w, style="hertz/history", precision=precision, skin=0.02, check=check)
    b = H.device_sim(w, style="hertz/history", precision=precision, skin=0.02, check=check)
    b.atom._native.set_launch_mode(True)
    for steps in (40, 17, 1, 23):
        a.run(steps)
        b.run(steps)
        assert a.last_run_device_resident and b.last_run_device_resident
        sa, sb = H.device_state(a), H.device_state(b)
        _assert_state_equal(sa, sb, ("x", "v", "omega", "f", "torque", "virial"))
        assert np.array_equal(sa["boxhi"], sb["boxhi"])
        assert a.neighbor.build_count == b.neighbor.build_count
    assert a.neighbor.build_count > 4


@pytest.mark.parametrize("precision", ["f64", "f32split"])
def test_fields_through_device_memory_equal_the_host_path(precision):
    """tmps_get_field_device reads every field as tmps_get_field does, and
    tmps_set_field_device leaves a system as tmps_set_field does: two
    systems given the same new positions, velocities and spins (one through
    host arrays, one through GPU tensors) run on bit for bit."""
    torch = pytest.importorskip("torch")
    from taichimps_amd import _native as NT
    w = dense(8, 1.0, seed=91)
    rng = np.random.default_rng(92)
    w["v"] = rng.normal(0.0, 1.0, size=w["v"].shape)
    a = H.device_sim(w, style="hertz/history", precision=precision, skin=0.02)
    b = H.device_sim(w, style="hertz/history", precision=precision, skin=0.02)
    for s in (a, b):
        s.run(20)
    n = a.atom.nlocal
    A, B = a.atom._native, b.atom._native
    for f in (NT.F_X, NT.F_V, NT.F_OMEGA, NT.F_F, NT.F_TORQUE, NT.F_RADIUS, NT.F_RMASS,
              NT.F_TAG, NT.F_VIRIAL, NT.F_X0):
        host = A.get_field(f, n)
        dev = A.get_field_device(f, n).cpu().numpy()
        assert dev.dtype == host.dtype and np.array_equal(host, dev), f
    x = A.get_field(NT.F_X, n) + rng.uniform(-0.01, 0.01, size=(n, 3))
    v = rng.normal(0.0, 1.0, size=(n, 3))
    om = rng.normal(0.0, 2.0, size=(n, 3))
    for f, vals in ((NT.F_X, x), (NT.F_V, v), (NT.F_OMEGA, om)):
        A.set_field(f, vals, n)
        B.set_field_device(f, torch.tensor(vals, dtype=torch.float64, device="cuda"), n)
        assert np.array_equal(A.get_field(f, n), B.get_field(f, n)), f
    for steps in (30, 11):
        a.run(steps)
        b.run(steps)
        sa, sb = H.device_state(a), H.device_state(b)
        _assert_state_equal(sa, sb, ("x", "v", "omega", "f", "torque", "virial"))
    assert a.neighbor.build_count == b.neighbor.build_count
