"""LAMMPS script driver (reference: input.py, data_reader.py, main.py).

CPU: expression evaluation (the reference's own variable test), the
argument grammar of pair_style / pair_coeff / fix deform / fix wall/gran with
their NULL shorthands, read_data, and the loud failure of commands outside
the path. GPU: scripts run through the parser give, bit for bit, the runs the
same systems give through the Python API (isotropic compression with the
servo, a restart resumed from a script, spheres on a floor wall).
"""

from __future__ import annotations

import numpy as np
import pytest

from taichimps_amd import input as I
from taichimps_amd.data_reader import read_data


def _parser(tmp_path, text=""):
    p = tmp_path / "in.test"
    p.write_text(text)
    return I.LAMMPSInputParser(p)


def test_variable_evaluation(tmp_path):
    """The reference's tests/test_input_parser.py case."""
    p = _parser(tmp_path, """
variable Poi equal 0.23
variable Gshear equal 29.1e9
variable kn equal 4.0*${Gshear}/(3*(1-${Poi}))
variable kt equal ${kn}*2*(1-${Poi})/(2-${Poi})
""")
    p.execute()
    assert abs(p.variables["Poi"] - 0.23) < 1e-6
    assert abs(p.variables["Gshear"] - 29.1e9) < 1e-6
    kn = 4.0 * 29.1e9 / (3.0 * (1.0 - 0.23))
    assert abs(p.variables["kn"] - kn) < 1e-3
    assert abs(p.variables["kt"] - kn * 2 * (1 - 0.23) / (2 - 0.23)) < 1e-3


def test_expressions_comments_and_continuations(tmp_path):
    p = _parser(tmp_path, """
variable a equal 2^3      # a comment
variable b equal sqrt(v_a*2)+PI-PI &
                 +abs(-1)
variable s string hello
variable x index 7
variable c equal ${a}*$x
""")
    p.execute()
    assert p.variables["a"] == 8.0
    assert p.variables["b"] == pytest.approx(5.0)
    assert p.variables["s"] == "hello"
    assert p.variables["c"] == 56.0
    with pytest.raises(ValueError):
        p.evaluate_expression("__import__('os')")
    with pytest.raises(ValueError):
        p.evaluate_expression("undefined_name + 1")


def test_gran_settings_null_shorthands():
    ev = float
    s = I.parse_gran_settings("2e5 NULL 50 NULL 0.5 1".split(), ev)
    assert s["kt"] == 2e5 * 2.0 / 7.0 and s["gammat"] == 25.0 and not s["limit_damping"]
    s = I.parse_gran_settings("2e5 1e5 50 10 0.5 0 limit_damping".split(), ev)
    assert s["kt"] == 1e5 and s["dampflag"] == 0 and s["limit_damping"]
    with pytest.raises(ValueError):
        I.parse_gran_settings("2e5 1e5 50 10 0.5 0 bogus".split(), ev)
    with pytest.raises(ValueError):
        I.parse_gran_settings("2e5 1e5 50".split(), ev)


def test_pair_coeff_granular_grammar():
    ev = float
    s = I.parse_pair_coeff_granular(
        "* * hertz/material 1e7 0.5 0.3 tangential mindlin NULL 1.0 0.5 damping "
        "coeff_restitution".split(), ev)
    assert s == dict(normal="hertz/material", normal_coeffs=[1e7, 0.5, 0.3],
                     tangential="mindlin", tangential_coeffs=[None, 1.0, 0.5],
                     damping="coeff_restitution", limit_damping=False)
    s = I.parse_pair_coeff_granular("* * hooke 1e5 0.2 limit_damping".split(), ev)
    assert s["damping"] == "viscoelastic" and s["limit_damping"]
    with pytest.raises(ValueError):
        I.parse_pair_coeff_granular("* * dmt 1 2 3 4".split(), ev)
    with pytest.raises(NotImplementedError):
        I.parse_pair_coeff_granular("* * hooke 1e5 0.2 rolling sds 1 2 3".split(), ev)


def test_deform_grammar():
    s = I.parse_deform("1 x pressure 1.0 1e-4 y erate -1e-3 z final 0 10 couple none "
                       "remap x max/rate 0.01".split(), float)
    assert s["nevery"] == 1 and s["couple"] == "none" and s["remap"] == "x"
    assert s["max_rate"] == 0.01
    assert s["axes"]["x"] == {"style": "pressure", "ptarget": 1.0, "pgain": 1e-4}
    assert s["axes"]["y"] == {"style": "erate", "rate": -1e-3}
    assert s["axes"]["z"] == {"style": "final", "flo": 0.0, "fhi": 10.0}
    with pytest.raises(ValueError):
        I.parse_deform("1 x bogus 1".split(), float)
    with pytest.raises(NotImplementedError):
        I.parse_deform("1 xy erate 0.1".split(), float)


def test_wall_gran_grammar():
    lo = (0.0, -1.0, 2.0)
    w = I.parse_wall_gran("hooke/history 1e5 NULL 10 NULL 0.5 1 zplane 0.5 NULL".split(), float, lo)
    assert len(w) == 1
    assert w[0]["wall_axis"] == 2 and w[0]["wall_side"] == -1 and w[0]["wall_coord"] == 0.5
    assert w[0]["kt"] == 1e5 * 2.0 / 7.0 and w[0]["gammat"] == 5.0 and w[0]["history"]
    w = I.parse_wall_gran("hooke 1e5 2e4 10 5 0.5 1 xplane -1 4".split(), float, lo)
    assert [(x["wall_side"], x["wall_coord"]) for x in w] == [(-1, -1.0), (1, 4.0)]
    assert not w[0]["history"]
    w = I.parse_wall_gran("hooke 1e5 2e4 10 5 0.5 1 yplane NULL NULL".split(), float, lo)
    assert w[0]["wall_coord"] == -1.0  # no bound: the box's lower face
    w = I.parse_wall_gran("hertz/history 1e5 NULL 10 NULL 0.5 1 zcylinder 3 1 2".split(),
                          float, lo)
    assert w == [dict(cylinder=True, fstyle="hertz/history", kn=1e5, gamman=10.0, kt=None,
                      gammat=None, xmu=0.5, dampflag=1, axis=2, radius=3.0, c1=1.0, c2=2.0,
                      side="in")]
    reg = {"c": dict(axis=0, c1=1.0, c2=2.0, radius=4.0, axis_lo=0.0, axis_hi=None,
                     side="out")}
    w = I.parse_wall_gran("hooke 1e5 2e4 10 5 0.5 1 region c".split(), float, lo, reg)
    assert w[0]["cylinder"] and w[0]["side"] == "out" and w[0]["axis_lo"] == 0.0
    with pytest.raises(NotImplementedError):
        I.parse_wall_gran("granular hooke 1e5 0.2 zplane 0 NULL".split(), float, lo)


def _write_data(path, w):
    r = lambda v: repr(float(v))  # noqa: E731  (shortest round-trip text)
    n = len(w["x"])
    lines = ["LAMMPS data file", "", f"{n} atoms", "1 atom types", ""]
    for k, a in enumerate("xyz"):
        lines.append(f"{r(w['boxlo'][k])} {r(w['boxhi'][k])} {a}lo {a}hi")
    lines += ["", "Atoms # sphere", ""]
    for i in range(n):
        x = w["x"][i]
        lines.append(f"{int(w['tag'][i])} 1 {r(2.0 * w['radius'][i])} {r(w['density'])} "
                     f"{r(x[0])} {r(x[1])} {r(x[2])}")
    lines += ["", "Velocities", ""]
    for i in range(n):
        v, o = w["v"][i], w["omega"][i]
        lines.append(f"{int(w['tag'][i])} {r(v[0])} {r(v[1])} {r(v[2])} "
                     f"{r(o[0])} {r(o[1])} {r(o[2])}")
    path.write_text("\n".join(lines) + "\n")


def test_read_data_sphere(tmp_path):
    from taichimps_amd import workloads
    w = workloads.lattice(3, 3, 3, 1.1, 4, "uniform")
    w["omega"] = np.random.default_rng(1).normal(size=w["x"].shape)
    _write_data(tmp_path / "d.data", w)
    d = read_data(tmp_path / "d.data")
    assert d.natoms == 27 and d.natom_types == 1
    assert d.boxlo == list(w["boxlo"]) and d.boxhi == list(w["boxhi"])
    assert np.array_equal(d.x, w["x"]) and np.array_equal(d.radius, 0.5 * (2.0 * w["radius"]))
    assert np.array_equal(d.v, w["v"]) and np.array_equal(d.omega, w["omega"])
    assert np.array_equal(d.tag, w["tag"]) and np.all(d.density == 1.0)
    with pytest.raises(FileNotFoundError):
        read_data(tmp_path / "missing.data")


def test_commands_outside_the_path_fail_loudly(tmp_path):
    p = _parser(tmp_path)
    for cmd in ("dump", "create_atoms", "velocity"):
        with pytest.raises(NotImplementedError):
            p.run_command(cmd, ["1", "all"])
    with pytest.raises(NotImplementedError):
        p.run_command("pair_style", ["lj/cut", "2.5"])
    with pytest.raises(ValueError):
        p.run_command("compute", ["9", "all", "rdf", "100"])
    with pytest.raises(ValueError):
        p.run_command("atom_style", ["atomic"])


def test_reduce_modes():
    v = np.array([1.0, -2.0, 3.0])
    got = {m: I._reduce(m, v) for m in I.REDUCE_MODES}
    assert got == dict(sum=2.0, ave=2.0 / 3.0, min=-2.0, max=3.0, sumsq=14.0, avesq=14.0 / 3.0,
                       sumabs=6.0, aveabs=2.0, maxabs=3.0)
    assert I._reduce("ave", np.zeros(0)) == 0.0


def test_cli_without_a_script_prints_help(capsys):
    from taichimps_amd import main as M
    assert M.main([]) == 0
    assert "LAMMPS" in capsys.readouterr().out


def _print_parser(tmp_path, text):
    from taichimps_amd.domain import Domain
    p = _parser(tmp_path, text)
    p.domain = Domain([0.0, 0.0, 0.0], [2.0, 3.0, 4.0], ("p", "p", "p"))
    p.execute()
    return p


def test_fix_print_keeps_its_text_and_reads_its_options(tmp_path):
    """The text is substituted when it prints, the other words when read
    (input.py:788-794, 1170-1204); a file gets the default header, a title,
    or none for a blank title (print.py:39-47)."""
    p = _print_parser(tmp_path, """
variable n equal 5
variable w equal 2*v_n
fix 1 all print ${n} "step ${step} w $w lx ${lx}" file out1.txt screen no
fix 2 all print 3 "a ${n}" file out2.txt title "# step a"
fix 3 all print 7 "b" file out3.txt title ""
fix 4 all print 2 "c" screen yes
""")
    f1, f2, f3, f4 = (p.fixes[k] for k in "1234")
    assert isinstance(f1, I.FixPrint) and f1.nevery == 5
    assert f1.template_str == "step ${step} w $w lx ${lx}"
    assert f1.filepath == tmp_path / "out1.txt" and not f1.screen
    assert f4.screen and f4.filepath is None
    for f in (f1, f2, f3):
        f.close()
    assert (tmp_path / "out1.txt").read_text() == "# Fix print output for fix 1\n"
    assert (tmp_path / "out2.txt").read_text() == "# step a\n"
    assert (tmp_path / "out3.txt").read_text() == ""


def test_fix_print_writes_on_absolute_multiples(tmp_path):
    """A multiple of N on the absolute timestep (print.py:49-66): a run
    resumed at step 1000003 with N = 4 prints at 1000004, 1000008."""
    out = []
    f = I.FixPrint(None, 4, "t", lambda t: out.append(t) or f"row {len(out)}",
                   filepath=tmp_path / "o.txt", fix_id="p")
    for step in range(1000003, 1000010):
        f.timestep = step
        f.end_of_step(None, 1e-4)
    # the device run's cut: 3 steps that end on a multiple print once
    f.timestep = 1000012
    f.advance(3)
    f.close()
    assert f.step_count == 10 and f.lines == ["row 1", "row 2", "row 3"]
    assert (tmp_path / "o.txt").read_text().splitlines() == [
        "# Fix print output for fix p", "row 1", "row 2", "row 3"]


def test_string_template_substitution(tmp_path):
    """${name} and $name (several characters, as the reference reads them),
    equal-style variables evaluated now with %.15g, stored strings as they
    are, unknown names as nothing, state keywords (input.py:115-198)."""
    p = _print_parser(tmp_path, """
variable n equal 1000000
variable third equal 1/3
variable tag string hello
variable s equal 2*step+lx
variable bad equal 1/0
variable Vs equal 6.0
""")
    p.start_timestep = 7
    ev = p.evaluate_string_template
    assert ev("${n} $n ${third}") == "1000000 1000000 0.333333333333333"
    assert ev("$tag-${tag}") == "hello-hello"
    assert ev("[${nope}]") == "[]"
    assert ev("${step} ${lx} ${ly} ${vol} ${s}") == "7 2 3 24 16"
    assert ev("${time}") == f"{7 * p.dt:.15g}"
    with pytest.raises(ValueError, match="bad"):
        ev("${bad}")
    # the void ratio from the stored solid volume (input.py:216-221)
    assert ev("${e}") == f"{(24.0 - 6.0) / 6.0:.15g}"
    # keywords are also names in expressions and in immediate ${} text
    assert p.evaluate_expression("lx*ly*lz") == 24.0
    assert p.evaluate_expression("${vol}/2") == 12.0


# ------------------------------------------------------------------ GPU

ISO = """
units           si
atom_style      sphere
boundary        p p p
read_data       lattice.data
pair_style      gran/hertz/history 2e5 NULL 50 NULL 0.5 1
pair_coeff      * *
neighbor        0.1 bin
neigh_modify    delay 0 every 1 check yes
timestep        1e-4
fix             1 all nve/sphere
fix             2 all deform/pressure 1 x pressure 1.0 1e-4 y pressure 1.0 1e-4 &
                z pressure 1.0 1e-4 couple xyz
compute         cn all contact/atom
compute         cnave all reduce ave c_cn
compute         p all pressure NULL pair
compute         s all stress/atom NULL
compute         sxx all reduce sum c_s[1]
compute         fab all fabric
thermo          1
thermo_style    custom step atoms press c_cnave c_p[1]
run             100
run             50
"""


def _api_sim(w, pair="hertz/history", deform=True):
    import simhelpers as H
    # the NULL shorthands as the script derives them: Kt = Kn * 2 / 7
    return H.device_sim(w, style=pair, servo=deform, kt=2e5 * 2.0 / 7.0, gammat=0.5 * 50.0)


@pytest.mark.gpu
def test_isotropic_script_equals_the_api_run_bitwise(tmp_path):
    from taichimps_amd import workloads
    w = workloads.lattice(8, 8, 8, 1.1, 1001, "uniform")
    _write_data(tmp_path / "lattice.data", w)
    (tmp_path / "in.iso").write_text(ISO)
    p = I.LAMMPSInputParser(tmp_path / "in.iso")
    p.execute()
    sim = _api_sim(w)
    sim.run(100)
    sim.run(50)
    n = len(w["x"])
    for key in ("x", "v", "omega", "f"):
        assert np.array_equal(getattr(p.atom, key).to_numpy()[:n],
                              getattr(sim.atom, key).to_numpy()[:n]), key
    assert np.array_equal(p.domain.boxhi, sim.domain.boxhi)
    assert p.simulation.timestep == 150 and p.simulation.last_run_device_resident
    assert len(p.thermo_lines) == 2
    step, atoms, press, cnave, pxx = p.thermo_lines[-1].split()
    assert int(step) == 150 and int(atoms) == n
    assert float(cnave) == pytest.approx(p.compute_value("cnave"))
    # per-atom stress summed by compute reduce is minus the pair pressure times volume
    vol = float(np.prod(np.asarray(p.domain.boxhi) - np.asarray(p.domain.boxlo)))
    assert p.compute_value("sxx") == pytest.approx(-p.compute_value("p", 1) * vol, rel=1e-10)
    fab = [p.compute_value("fab", k) for k in (1, 2, 3)]
    assert sum(fab) == pytest.approx(1.0, abs=1e-12)


@pytest.mark.gpu
def test_restart_written_and_read_by_scripts_resumes_bitwise(tmp_path):
    from taichimps_amd import workloads
    w = workloads.lattice(6, 6, 6, 0.98, 5, "uniform")
    _write_data(tmp_path / "lattice.data", w)
    head = """
atom_style sphere
boundary p p p
read_data lattice.data
pair_style granular
pair_coeff * * hertz/material 1e7 0.5 0.3 tangential mindlin NULL 1.0 0.5 damping coeff_restitution
neighbor 0.1 bin
timestep 1e-4
fix 1 all nve/sphere
"""
    (tmp_path / "in.cont").write_text(head + "run 80\nrun 70\n")
    (tmp_path / "in.first").write_text(head + "run 80\nwrite_restart mid.restart\n")
    (tmp_path / "in.second").write_text("""
atom_style sphere
read_restart mid.restart
neighbor 0.1 bin
timestep 1e-4
fix 1 all nve/sphere
run 70
""")
    cont = I.LAMMPSInputParser(tmp_path / "in.cont")
    cont.execute()
    I.LAMMPSInputParser(tmp_path / "in.first").execute()
    second = I.LAMMPSInputParser(tmp_path / "in.second")
    second.execute()
    n = len(w["x"])
    assert second.simulation.timestep == 150
    for key in ("x", "v", "omega", "f", "torque"):
        assert np.array_equal(getattr(cont.atom, key).to_numpy()[:n],
                              getattr(second.atom, key).to_numpy()[:n]), key


@pytest.mark.gpu
def test_floor_wall_script_equals_the_api_run_bitwise(tmp_path):
    import taichimps_amd as tm
    from taichimps_amd import workloads
    w = workloads.lattice(5, 5, 3, 1.1, 13, "uniform")
    w["v"][:, 2] -= 2.0
    _write_data(tmp_path / "lattice.data", w)
    floor = float(w["boxlo"][2]) + 0.08
    (tmp_path / "in.wall").write_text(f"""
atom_style sphere
boundary p p f
read_data lattice.data
pair_style gran/hooke/history 2e5 NULL 50 NULL 0.5 1
pair_coeff * *
neighbor 0.1 bin
timestep 1e-4
fix 1 all nve/sphere
fix 2 all wall/gran hooke/history 2e5 NULL 50 NULL 0.5 1 zplane {floor!r} NULL
run 60
""")
    p = I.LAMMPSInputParser(tmp_path / "in.wall")
    p.execute()
    n = len(w["x"])
    dom = tm.Domain(w["boxlo"], w["boxhi"], ("p", "p", "f"))
    atom = tm.AtomSystem(max_atoms=n)
    atom.add_particles(x=w["x"], radius=w["radius"], density=w["density"], v=w["v"],
                       omega=w["omega"], tag=w["tag"])
    nl = tm.NeighborList(dom, skin=0.1, max_atoms=n, max_neighbors_per_atom=64)
    c = dict(kn=2e5, gamman=50.0, kt=2e5 * 2.0 / 7.0, gammat=25.0, xmu=0.5)
    sim = tm.Simulation(dom, atom, nl, tm.ContactHistory(n, 64),
                        tm.GranHookeHistory(dom, c["kn"], c["gamman"], c["kt"], c["gammat"],
                                            c["xmu"]), dt=1e-4)
    sim.add_fix(tm.FixNVESphere(dom))
    sim.add_fix(tm.FixWallGran(dom, 2, -1, floor, c["kn"], c["gamman"], c["kt"], c["gammat"],
                               c["xmu"], max_atoms=n))
    sim.run(60)
    for key in ("x", "v", "omega"):
        assert np.array_equal(getattr(p.atom, key).to_numpy()[:n],
                              getattr(atom, key).to_numpy()[:n]), key


@pytest.mark.gpu
def test_cylinder_region_wall_script_equals_the_api_run_bitwise(tmp_path):
    import taichimps_amd as tm
    from taichimps_amd import workloads
    w = workloads.lattice(4, 4, 4, 1.1, 31, "uniform")
    w["v"][:, 0] += 1.0
    _write_data(tmp_path / "lattice.data", w)
    lo, hi = w["boxlo"], w["boxhi"]
    cx, cy = float(0.5 * (lo[0] + hi[0])), float(0.5 * (lo[1] + hi[1]))
    (tmp_path / "in.cyl").write_text(f"""
atom_style sphere
boundary f f f
read_data lattice.data
pair_style gran/hooke/history 2e5 NULL 50 NULL 0.5 1
pair_coeff * *
neighbor 0.1 bin
timestep 1e-4
region can cylinder z {cx!r} {cy!r} 2.6 INF INF side in
fix 1 all nve/sphere
fix 2 all wall/gran/region hertz/history 2e5 NULL 50 NULL 0.5 1 region can
run 80
""")
    p = I.LAMMPSInputParser(tmp_path / "in.cyl")
    p.execute()
    n = len(w["x"])
    dom = tm.Domain(lo, hi, ("f", "f", "f"))
    atom = tm.AtomSystem(max_atoms=n)
    atom.add_particles(x=w["x"], radius=w["radius"], density=w["density"], v=w["v"],
                       omega=w["omega"], tag=w["tag"])
    nl = tm.NeighborList(dom, skin=0.1, max_atoms=n, max_neighbors_per_atom=64)
    c = dict(kn=2e5, gamman=50.0, kt=2e5 * 2.0 / 7.0, gammat=25.0, xmu=0.5)
    sim = tm.Simulation(dom, atom, nl, tm.ContactHistory(n, 64),
                        tm.GranHookeHistory(dom, c["kn"], c["gamman"], c["kt"], c["gammat"],
                                            c["xmu"]), dt=1e-4)
    sim.add_fix(tm.FixNVESphere(dom))
    sim.add_fix(tm.FixWallGranRegion(dom, axis="z", c1=cx, c2=cy, radius=2.6, side="in",
                                     fstyle="hertz/history", kn=2e5, gamman=50.0, kt=None,
                                     gammat=None, xmu=0.5, max_atoms=n))
    sim.run(80)
    for key in ("x", "v", "omega"):
        assert np.array_equal(getattr(p.atom, key).to_numpy()[:n],
                              getattr(atom, key).to_numpy()[:n]), key
    wall = p.fixes["2"]
    assert isinstance(wall, tm.FixWallGranRegion) and wall.touch.to_numpy()[:n].sum() > 0


@pytest.mark.gpu
def test_periodic_restart_files_leave_the_run_unchanged(tmp_path):
    """`restart 30 snap.*.restart` cuts the device run at steps 30, 60, 90
    (one fix setup, no Verlet::setup between the pieces): the trajectory is
    the unbroken run's bit for bit, and each file holds its step."""
    from taichimps_amd import restart as RS
    from taichimps_amd import workloads
    w = workloads.lattice(8, 8, 8, 1.1, 1001, "uniform")
    _write_data(tmp_path / "lattice.data", w)
    script = ISO.replace("run             100\nrun             50\n",
                         "restart         30 snap.*.restart\nrun             100\n")
    assert script != ISO
    (tmp_path / "in.snap").write_text(script)
    p = I.LAMMPSInputParser(tmp_path / "in.snap")
    p.execute()
    sim = _api_sim(w)
    sim.run(100)
    n = len(w["x"])
    for key in ("x", "v", "omega", "f"):
        assert np.array_equal(getattr(p.atom, key).to_numpy()[:n],
                              getattr(sim.atom, key).to_numpy()[:n]), key
    assert np.array_equal(p.domain.boxhi, sim.domain.boxhi)
    for step in (30, 60, 90):
        d = RS.read_restart(tmp_path / f"snap.{step}.restart")
        assert d.timestep == step and d.natoms == n
        # the file holds the state of step `step`: an unbroken run of that
        # many steps, written by write_restart, has the same atoms, box and
        # contact history records
        ref_script = ISO.replace("run             100\nrun             50\n",
                                 f"run             {step}\nwrite_restart   ref.{step}.restart\n")
        (tmp_path / f"in.ref{step}").write_text(ref_script)
        I.LAMMPSInputParser(tmp_path / f"in.ref{step}").execute()
        r = RS.read_restart(tmp_path / f"ref.{step}.restart")
        assert r.timestep == step
        for key in ("tag", "x", "v", "omega", "radius", "rmass"):
            assert np.array_equal(getattr(d, key), getattr(r, key)), (step, key)
        assert d.boxlo == r.boxlo and d.boxhi == r.boxhi
        assert len(d.fix_extra) == len(r.fix_extra) == n
        assert sum(len(e) for e in d.fix_extra) > 0
        for a, b in zip(d.fix_extra, r.fix_extra):
            assert np.array_equal(a, b), step
    assert not (tmp_path / "snap.100.restart").exists()


PRINT = """
units           si
atom_style      sphere
boundary        p p p
read_data       lattice.data
pair_style      gran/hertz/history 2e5 NULL 50 NULL 0.5 1
pair_coeff      * *
neighbor        0.1 bin
neigh_modify    delay 0 every 1 check yes
timestep        1e-4
fix             1 all nve/sphere
fix             2 all deform/pressure 1 x pressure 1.0 1e-4 y pressure 1.0 1e-4 &
                z pressure 1.0 1e-4 couple xyz
compute         1 all pressure thermo_temp
compute         cn all contact/atom
compute         cnave all reduce ave c_cn
variable        pxx equal c_1[1]
variable        pyy equal c_1[2]
variable        pzz equal c_1[3]
variable        pxy equal c_1[4]
variable        p equal (v_pxx+v_pyy+v_pzz)/3
variable        lx equal lx
variable        cn equal c_cnave
variable        Vs equal {vs!r}
variable        e equal (vol-v_Vs)/v_Vs
variable        KEtra equal ke_trans
variable        KErot equal ke_rot
fix             out1 all print 20 "${{step}} ${{pxx}} ${{pyy}} ${{pzz}} ${{pxy}} ${{p}}" &
                file mean_stress.txt title "# step pxx pyy pzz pxy p"
fix             out2 all print 25 "${{step}} ${{lx}} ${{ly}} ${{lz}} ${{e}} ${{cn}}" file box.txt
fix             out3 all print 50 "${{step}} ${{KEtra}} ${{KErot}}" file energy.txt screen no
run             100
"""


def _rows(path):
    return [line.split() for line in path.read_text().splitlines() if not line.startswith("#")]


@pytest.mark.gpu
def test_fix_print_output_equals_the_oracle_run(tmp_path):
    """A script with three fix print outputs (periods 20, 25, 50) runs
    device-resident, cut at the print steps, and every printed value is the
    oracle's at that step: pressure tensor, mean pressure, box lengths, void
    ratio and mean contact number as the same %.15g text (the FP64 path is
    bitwise), kinetic energies within 1e-12 (the oracle's are numpy sums, the
    device's a pairwise tree)."""
    import simhelpers as H
    from taichimps_amd import workloads
    w = workloads.lattice(8, 8, 8, 1.1, 1001, "uniform")
    _write_data(tmp_path / "lattice.data", w)
    vs = float(np.sum(4.0 / 3.0 * np.pi * w["radius"] ** 3))
    (tmp_path / "in.print").write_text(PRINT.format(vs=vs))
    p = I.LAMMPSInputParser(tmp_path / "in.print")
    p.execute()
    assert p.simulation.timestep == 100 and p.simulation.last_run_device_resident

    o = H.oracle_sim(w, style="hertz/history", kt=2e5 * 2.0 / 7.0, gammat=0.5 * 50.0)
    dt = 1e-4
    o.servo_setup(100, dt)
    assert o.init_simulation(dt) == 0
    m = 4.0 / 3.0 * np.pi * w["radius"] ** 3 * w["density"]
    want = {"mean_stress.txt": [], "box.txt": [], "energy.txt": []}
    g = (lambda v: f"{v:.15g}")
    for step in range(1, 101):
        assert o.step(dt) == 0
        lo, hi, _ = o.box()
        prd = hi - lo
        vol = float(prd[0] * prd[1] * prd[2])
        if step % 20 == 0:
            t = o.virial(keflag=True) / vol
            want["mean_stress.txt"].append([str(step)] + [g(v) for v in t[:4]]
                                           + [g((t[0] + t[1] + t[2]) / 3)])
        if step % 25 == 0:
            cn = float(np.mean(o.contact_atom().astype(np.float64)))
            want["box.txt"].append([str(step)] + [g(v) for v in prd]
                                   + [g((vol - vs) / vs), g(cn)])
        if step % 50 == 0:
            a = o.atoms()
            ket = 0.5 * np.sum(m * np.sum(a["v"] ** 2, axis=1))
            ker = 0.5 * np.sum(0.4 * m * w["radius"] ** 2 * np.sum(a["omega"] ** 2, axis=1))
            want["energy.txt"].append([step, ket, ker])
    assert (tmp_path / "mean_stress.txt").read_text().startswith("# step pxx pyy pzz pxy p\n")
    assert (tmp_path / "box.txt").read_text().startswith("# Fix print output for fix out2\n")
    assert _rows(tmp_path / "mean_stress.txt") == want["mean_stress.txt"]
    assert _rows(tmp_path / "box.txt") == want["box.txt"]
    got = _rows(tmp_path / "energy.txt")
    assert [int(r[0]) for r in got] == [50, 100]
    for r, (step, ket, ker) in zip(got, want["energy.txt"]):
        assert float(r[1]) == pytest.approx(ket, rel=1e-12)
        assert float(r[2]) == pytest.approx(ker, rel=1e-12)
    # the cuts leave the trajectory the oracle's
    dev, orc = H.device_state(p.simulation), H.oracle_state(o)
    for key in ("x", "v", "omega", "boxlo", "boxhi"):
        assert np.array_equal(dev[key], orc[key]), key
