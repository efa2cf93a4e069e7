"""Benchmark of the DEM step on one MI355X.

Workload: isotropic compression of 1,000,000 spheres (100^3 jittered lattice,
a = 1.1, radii U[0.45, 0.55], seed 1002), gran/hertz/history kn = 2e5,
kt = 2/7 kn, gamman = 50, gammat = 25, xmu = 0.5, dt = 1e-4, skin 0.1,
nve/sphere + deform/pressure isotropic P = 1.0 (pgain 1e-4, couple xyz,
remap x). Synthetic inputs from taichimps_amd.workloads.

For each precision (FP64 plain, FP32-split) the script runs `--warmup` steps
(the run's setup force and first neighbor build included), captures the
device graphs of the timed run (no step runs), then times exactly `--steps`
device-resident steps (Simulation.run with `pre no`) with HIP events recorded
on the system's stream around the run, device synchronised on both sides, on
`--repeats` fresh systems (the median run is reported; the runs are bitwise
equal). `value` is the FP32-split throughput in atom-steps/s; the FP64 figure
and their ratio ride along. A further pass of `--prof-steps` steps with
per-kernel HIP events gives the step kernel's average duration for the
roofline. The `c2_200` block runs both precisions 200 steps from the start
(rebuilds included and counted) on two fresh systems each, reports the second
(the first run in the process also pays one-time costs, kept beside it), and
compares their mean pressure with the CPU oracle's after those 200 steps. The CPU baseline times the oracle's OpenMP
build on the host's threads on the same workload and on the 8,000-sphere
hooke/history case; its serial build checks the FP64 device positions bit for
bit. The walls (1M spheres between a sheared floor and a ceiling), adjoint
(32k spheres), ensemble (256 packings) and 4M-sphere triaxial blocks follow
in the same JSON line (each can be switched off: --no-walls, --no-adjoint,
--no-ensemble, --no-triaxial).

    python bench.py [--gpus 1] [--steps 200] [--warmup 10]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def build_sim(w, precision):
    from taichimps_amd import workloads
    return workloads.build_isotropic(w, precision)


# Algorithmic bytes per atom-step of the one-kernel step (own-atom streams,
# list slots and touching-slot history; neighbour gathers are cache traffic
# and not counted): B = B_atom + 6 n_slot + B_hist n_touch, with
#   FP64   B_atom = x r+w 48 + v r+w 48 + omega r+w 48 + radius, mass 16
#                   + x0 r 24 + list offset 4 = 188,   B_hist = shear r+w 48
#   split  B_atom = offset r+w 24 + cell 4 + v r+w 24 + omega r+w 24
#                   + radius, mass 8 + x0 offset 12 + list offset 4 = 100,
#                   B_hist = 24
# and 6 B per slot (4 B index, 1 B touch flag read and written).
STEP_MODEL = {"f64": (188.0, 6.0, 48.0), "f32split": (100.0, 6.0, 24.0)}


def model_bytes(key, n_slot, n_touch):
    base, per_slot, per_touch = key
    return base + per_slot * n_slot + per_touch * n_touch


def build_triaxial(w, precision):
    from taichimps_amd import workloads
    return workloads.build_triaxial(w, precision)


def triaxial_c3(steps=1000, warmup=10):
    """Triaxial test on 4,000,000 polydisperse spheres (160 x 125 x 200
    lattice, a = 1.3, log-normal radii, seed 1003; build_triaxial), FP32-split
    and FP64, `steps` timed steps after `warmup`. Rolling resistance is not
    part of it: the reference's PairGranular raises on rolling models."""
    from taichimps_amd import workloads
    w = workloads.case("triaxial_4m")
    out = {}
    for prec in ("f32split", "f64"):
        r = run_precision(w, prec, steps, warmup, 50, builder=build_triaxial)
        r.pop("x")
        out[prec] = r
        print(f"# triaxial {prec}: {r['ms_per_step']:.4f} ms/step, "
              f"{r['atom_steps_per_s']:.4e} atom-steps/s, kernels {r['kernel_ms']}, "
              f"rebuilds {r['rebuilds']}", file=sys.stderr, flush=True)
    ps, pd = np.array(out["f32split"]["pressure"]), np.array(out["f64"]["pressure"])
    return {
        "workload": "triaxial, 4,000,000 polydisperse spheres, PairGranular hertz/material + "
                    "viscoelastic + mindlin, damping/cundall 0.1, x/y pressure 1.0, z erate "
                    f"-1e-3, {steps} steps",
        "atoms": out["f32split"]["n"],
        "f32split": out["f32split"], "f64": out["f64"],
        "speedup_fp32split_vs_fp64": out["f32split"]["atom_steps_per_s"]
        / out["f64"]["atom_steps_per_s"],
        "pressure_xyz_rel_err_vs_fp64": (np.abs(ps[:3] - pd[:3]) / np.abs(pd[:3])).tolist(),
    }


def wall_run(steps=100, warmup=10):
    """Walls inside the device-resident step: the 1M-sphere lattice of the
    headline workload with z bounded by a sheared floor wall and a ceiling
    wall on the box faces (x/y periodic; about 970 spheres touch each wall at
    the start), gran/hertz/history, nve/sphere and damping/cundall 0.1, the
    spheres thrown down at 0.5, FP32-split and FP64. The same run without the
    walls (z periodic) is timed beside it."""
    from taichimps_amd import workloads
    w = workloads.case("isotropic_hertz_1m")
    n = len(w["x"])
    out = {"workload": f"{n} spheres, z between a sheared floor and a ceiling wall, "
                       f"gran/hertz/history, nve/sphere, damping/cundall 0.1, {steps} steps "
                       f"after {warmup}", "atoms": n}
    for prec in ("f32split", "f64"):
        row = {}
        for walls in (True, False):
            sim = wall_sim(w, prec, walls)
            sim.run(warmup)
            sim.run_pre = False
            sim.prepare_run(steps)
            b0 = sim.atom._native.counters()["build_count"]
            el = timed_run(sim, steps)
            key = "walls" if walls else "no_walls"
            row[key] = {"ms_per_step": el / steps * 1e3, "atom_steps_per_s": n * steps / el,
                        "device_resident": bool(sim.last_run_device_resident),
                        "list_builds": sim.atom._native.counters()["build_count"] - b0}
            if walls:
                row[key]["floor_touching_atoms"] = int(
                    sim.fixes[-2].touch.to_numpy()[:n].sum())
        row["wall_cost_ms_per_step"] = row["walls"]["ms_per_step"] - row["no_walls"]["ms_per_step"]
        out[prec] = row
        print(f"# walls {prec}: {row}", file=sys.stderr, flush=True)
    return out


def wall_sim(w, precision, walls):
    """The wall run's system (see wall_run); walls=False keeps z periodic."""
    import taichimps_amd as tm
    from taichimps_amd import workloads
    c = workloads.ISOTROPIC
    n = len(w["x"])
    v = w["v"].copy()
    v[:, 2] -= 0.5
    dom = tm.Domain(w["boxlo"], w["boxhi"], ("p", "p", "f" if walls else "p"))
    atom = tm.AtomSystem(max_atoms=n, float_type=precision)
    atom.add_particles(x=w["x"], radius=w["radius"], density=w["density"], v=v,
                       omega=w["omega"], tag=w["tag"])
    nl = tm.NeighborList(dom, skin=c["skin"], max_atoms=n, max_neighbors_per_atom=64)
    pair = tm.GranHertzHistory(dom, c["kn"], c["gamman"], c["kt"], c["gammat"], c["xmu"])
    sim = tm.Simulation(dom, atom, nl, tm.ContactHistory(n, 64), pair, dt=c["dt"])
    sim.add_fix(tm.FixNVESphere(dom))
    sim.add_fix(tm.FixDampingCundall(dom, 0.1, 0.1))
    if walls:
        lo, hi = float(w["boxlo"][2]), float(w["boxhi"][2])
        sim.add_fix(tm.FixWallGran(dom, 2, -1, lo, c["kn"], c["gamman"], c["kt"],
                                   c["gammat"], c["xmu"], shear_axis=0, shear_vel=0.2,
                                   max_atoms=n))
        sim.add_fix(tm.FixWallGran(dom, 2, 1, hi, c["kn"], c["gamman"], c["kt"],
                                   c["gammat"], c["xmu"], max_atoms=n))
    return sim


def timed_run(sim, steps):
    """Wall time (s) of sim.run(steps) on the device's clock: HIP events
    recorded on the system's stream (a torch stream handed to the system)
    before and after the run, the device synchronised on both sides."""
    import torch
    N = sim.atom._native
    stream = torch.cuda.Stream()
    N.set_stream(stream.cuda_stream)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    N.sync()
    start.record(stream)
    sim.run(steps)
    stop.record(stream)
    N.sync()
    torch.cuda.synchronize()
    N.set_stream(None)
    return start.elapsed_time(stop) * 1e-3


def run_precision(w, precision, steps, warmup, prof_steps, direct=False, builder=None, repeats=1):
    """`repeats` fresh systems, each with `warmup` steps then exactly `steps`
    timed steps (timed_run); the median timed run is reported (every run does
    the same work: the trajectories are bitwise equal)."""
    times = []
    p_first = None
    for _ in range(max(1, repeats)):
        sim = (builder or build_sim)(w, precision)
        N = sim.atom._native
        N.set_launch_mode(direct)
        n = sim.atom.nlocal
        t0 = time.perf_counter()
        sim.run(warmup)
        N.sync()
        t_setup = time.perf_counter() - t0
        sim.run_pre = False
        builds0 = sim.neighbor.build_count
        sim.prepare_run(steps)
        times.append(timed_run(sim, steps))
        p_run = N.last_pressure()
        if p_first is None:
            p_first = p_run
        elif not np.array_equal(p_first, p_run):
            raise RuntimeError("repeated runs of the same workload differ")
    elapsed = float(np.median(times))
    c = N.counters()
    rebuilds = sim.neighbor.build_count - builds0
    pressure = N.last_pressure()
    x_final = sim.atom.x.to_numpy()[:n]
    n_slot = c["nslots"] / n
    n_touch = c["ntouch"] / n
    # per-kernel HIP-event timings over the same number of steps, same stream
    ms, cnt = np.zeros(4), np.zeros(4, dtype=int)
    if prof_steps > 0:
        N.set_profiling(True)
        sim.run(prof_steps)
        ms, cnt = N.kernel_times()
        N.set_profiling(False)
    return {
        "n": n, "elapsed_s": elapsed, "ms_per_step": elapsed / steps * 1e3,
        "ms_per_step_runs": [t / steps * 1e3 for t in times],
        "atom_steps_per_s": n * steps / elapsed, "setup_s": t_setup, "rebuilds": rebuilds,
        "n_slot": n_slot, "n_touch": n_touch, "pressure": pressure.tolist(),
        "mean_pressure": float(pressure[:3].mean()),
        "kernel_ms": {"kick": float(ms[0]), "step": float(ms[1]), "tail": float(ms[2]),
                      "close": float(ms[3])},
        "kernel_counts": cnt.tolist(), "x": x_final,
    }


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_step_split_isotropic_hertz_1m_hbm_traffic.md")


def profiled_traffic():
    """HBM bytes per launch of the FP32-split step kernel as measured by the
    FETCH_SIZE / WRITE_SIZE counter passes recorded in the profile summary
    (each counter scaled by its plain-copy calibration factor), or None."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_calibration(mib):
    """A plain device copy of known size: the profiler's FETCH_SIZE / WRITE_SIZE
    for it, divided by the bytes it must move, is the counters' scale factor."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    # the same copy timed: what a streaming kernel reaches on this device
    reps = 20
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "copies": 5 + reps,
                      "copy_ms": ms, "copy_GBs_read_plus_write": 2 * n * 4 / (ms * 1e-3) / 1e9}))
    return 0


def cpu_baseline(w, warmup, steps, threaded=False, style="hertz/history"):
    """The CPU oracle on the same workload, with the same run split as the
    device (a `warmup` run with setup, then `steps` timed): single-threaded,
    or its OpenMP build on the host's threads (`threaded`)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import oracle as O
    from taichimps_amd.workloads import ISOTROPIC as C
    O.build()
    o = O.from_workload(w, threaded=threaded)
    o.set_neigh(C["skin"])
    o.set_pair(O.legacy_pair(style, kn=C["kn"], gamman=C["gamman"], kt=C["kt"],
                             gammat=C["gammat"], xmu=C["xmu"]))
    o.set_servo(O.servo_params(p_target=C["p_target"], p_gain=C["p_gain"]))
    assert o.run(warmup, C["dt"]) == 0
    t0 = time.perf_counter()
    assert o.run(steps, C["dt"], run_pre=False) == 0
    elapsed = time.perf_counter() - t0
    p = o.last_pressure_tensor()
    return {"atom_steps_per_s": len(w["x"]) * steps / elapsed, "elapsed_s": elapsed,
            "mean_pressure": float(p[:3].mean()), "x": o.atoms()["x"],
            "threads": O.num_threads(threaded)}


def c2_200(steps=200):
    """The 1M-sphere isotropic compression from its start: `steps` steps of
    each precision after the setup force (run with `pre yes`: Verlet::setup
    and the first neighbor build untimed, in a 0-step run), timed with
    timed_run, rebuilds counted, on two fresh systems (the second reported,
    the first's rate kept beside it); the final pressures are compared with
    the CPU oracle's in the CPU baseline leg."""
    from taichimps_amd import workloads
    w = workloads.case("isotropic_hertz_1m")
    out = {"workload": f"isotropic compression, 1M spheres, {steps} steps from the start",
           "note": "two fresh systems per precision, each timed over the same steps from the "
                   "start; the first run in the process also pays one-time costs of the "
                   "process (first_run_ms_per_step), the second is the figure"}
    for prec in ("f32split", "f64"):
        first = None
        for attempt in range(2):
            sim = build_sim(w, prec)
            sim.run(0)
            sim.run_pre = False
            builds0 = sim.neighbor.build_count
            sim.prepare_run(steps)
            elapsed = timed_run(sim, steps)
            if attempt == 0:
                first = elapsed
        p = sim.atom._native.last_pressure()
        n = sim.atom.nlocal
        out[prec] = {"atom_steps_per_s": n * steps / elapsed, "ms_per_step": elapsed / steps * 1e3,
                     "first_run_ms_per_step": first / steps * 1e3,
                     "rebuilds": sim.neighbor.build_count - builds0,
                     "pressure": p.tolist(), "mean_pressure": float(p[:3].mean())}
        if prec == "f64":
            out["f64_x"] = sim.atom.x.to_numpy()[:n]
        print(f"# c2_200 {prec}: {out[prec]}", file=sys.stderr, flush=True)
    out["speedup_fp32split_vs_fp64"] = (out["f32split"]["atom_steps_per_s"]
                                        / out["f64"]["atom_steps_per_s"])
    return out


def cpu_c2_200(steps=200):
    """The serial CPU oracle's run of the same 200 steps from the start (the
    checker of the c2_200 block)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import oracle as O
    from taichimps_amd import workloads
    from taichimps_amd.workloads import ISOTROPIC as C
    w = workloads.case("isotropic_hertz_1m")
    o = O.from_workload(w)
    o.set_neigh(C["skin"])
    o.set_pair(O.legacy_pair("hertz/history", kn=C["kn"], gamman=C["gamman"], kt=C["kt"],
                             gammat=C["gammat"], xmu=C["xmu"]))
    o.set_servo(O.servo_params(p_target=C["p_target"], p_gain=C["p_gain"]))
    assert o.run(0, C["dt"]) == 0
    t0 = time.perf_counter()
    assert o.run(steps, C["dt"], run_pre=False) == 0
    elapsed = time.perf_counter() - t0
    p = o.last_pressure_tensor()
    return {"mean_pressure": float(p[:3].mean()), "x": o.atoms()["x"], "elapsed_s": elapsed,
            "build_count": o.counters()["build_count"]}


def adjoint_c4(steps=2000, every=100, fd_rel=1e-6):
    """Adjoint calibration on 32,768 spheres (32^3 jittered lattice, a = 1.1,
    radii U[0.45, 0.55], seed 1004), gran/hertz/history, isotropic compression
    to P = 1.0 for `steps` steps, checkpoints every `every` steps, FP64.
    Targets come from a forward run with (kn, xmu, gamman) scaled by
    U[0.9, 1.1] drawn from default_rng([1004, 1]). The loss is
    (P - P*)^2 + (Z - Z*)^2 (P: final mean pressure, Z: coordination number);
    Z is piecewise constant, so its gradient is zero almost everywhere and the
    loss gradient is 2 (P - P*) dP/dtheta. The adjoint's dP/d(kn, xmu, gamman)
    is checked against central differences at relative step `fd_rel`."""
    from taichimps_amd import adjoint as ADJ
    from taichimps_amd import workloads
    from taichimps_amd.workloads import ISOTROPIC as C
    w = workloads.case("adjoint_32k")
    theta = np.array([C["kn"], C["kt"], C["gamman"], C["gammat"], C["xmu"]])
    f = np.random.default_rng([1004, 1]).uniform(0.9, 1.1, size=3)
    theta_t = theta.copy()
    theta_t[0] *= f[0]
    theta_t[4] *= f[1]
    theta_t[2] *= f[2]
    build = ADJ.isotropic_builder(w, style="hertz/history", skin=C["skin"], dt=C["dt"],
                                  p_target=C["p_target"], p_gain=C["p_gain"])
    mean_p = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0]) / 3.0

    def forward(th):
        sim = build(th, w["radius"])
        t0 = time.perf_counter()
        tensor, z = ADJ.run_forward(sim, steps, every)
        return sim, float(mean_p @ tensor), z, time.perf_counter() - t0

    _, p_t, z_t, _ = forward(theta_t)
    sim, p, z, t_fwd = forward(theta)
    loss = (p - p_t) ** 2 + (z - z_t) ** 2
    t0 = time.perf_counter()
    g_c, g_r, _ = ADJ.run_backward(sim, mean_p)
    t_bwd = time.perf_counter() - t0
    def central(th, k, rel, fwd):
        h = rel * th[k]
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        return (fwd(tp) - fwd(tm)) / (2.0 * h)

    sweep = (1e-5, 1e-6, 1e-7, 1e-8, 1e-9)
    checks = {}
    for k, name in ((0, "kn"), (4, "xmu"), (2, "gamman")):
        entry = {"adjoint": float(g_c[k])}
        for rel in sweep:
            fd = central(theta, k, rel, lambda th: forward(th)[1])
            entry[f"fd_{rel:g}"] = fd
            entry[f"rel_err_{rel:g}"] = abs(g_c[k] - fd) / abs(fd)
        entry["best_rel_err"] = min(entry[f"rel_err_{r:g}"] for r in sweep)
        checks[name] = entry
    # the same packing over a horizon short enough for central differences
    # to be converged at the prescribed relative step
    short = 200
    sim_s = build(theta, w["radius"])
    ADJ.run_forward(sim_s, short, every)
    g_s, _, _ = ADJ.run_backward(sim_s, mean_p)

    def fwd_short(th):
        t, _ = ADJ.run_forward(build(th, w["radius"]), short, every)
        return float(mean_p @ t)

    short_checks = {}
    for k, name in ((0, "kn"), (4, "xmu"), (2, "gamman")):
        fd = central(theta, k, fd_rel, fwd_short)
        short_checks[name] = {"adjoint": float(g_s[k]), f"fd_{fd_rel:g}": fd,
                              "rel_err": abs(g_s[k] - fd) / abs(fd)}
    return {
        "workload": "adjoint calibration, 32,768 spheres, gran/hertz/history, "
                    f"{steps} steps, checkpoints every {every}",
        "loss": loss, "P": p, "P_target": p_t, "Z": z, "Z_target": z_t,
        "dL_dtheta": dict(zip(("kn", "kt", "gamman", "gammat", "xmu"),
                              (2.0 * (p - p_t) * g_c).tolist())),
        "dP_dradius_norm": float(np.linalg.norm(g_r)),
        "fd_check_dP": checks, "fd_rel_step": fd_rel,
        "max_rel_err_fd_1e-6": max(c[f"rel_err_{fd_rel:g}"] for c in checks.values()),
        "max_best_rel_err_fd_sweep": max(c["best_rel_err"] for c in checks.values()),
        f"fd_check_dP_{short}_steps": short_checks,
        f"max_rel_err_fd_1e-6_{short}_steps": max(c["rel_err"] for c in short_checks.values()),
        "fd_note": "over 2000 steps the packing's collision sequence amplifies a 1e-6 "
                   "perturbation beyond the linear regime, while steps of 1e-8 and below "
                   "drown in the forward's rounding noise: the sweep shows central "
                   "differences approaching the adjoint and then scattering; over 200 "
                   "steps of the same packing they converge at 1e-6",
        "forward_s": t_fwd, "backward_s": t_bwd, "rebuilds": sim.neighbor.build_count,
        "note": "Z has zero gradient almost everywhere (piecewise constant)",
        "_check_inputs": (w, theta, steps, float(p), g_c, g_r, sim.neighbor.build_count),
    }


def cpu_adjoint_check(block):
    """The checker of the adjoint block: the tangent-linear restatement
    (oracle/tlm.cpp) differentiates the same 2,000-step run forward on the
    host's threads, with the forward's contact onsets, slips and rebuilds and
    no step size, so a change of the collision sequence cannot enter it. Its
    FP64 values must be the device forward's bit for bit; its derivatives
    d<P>/d(kn, kt, gamman, gammat, xmu) and along a random radius direction
    are compared with the adjoint's."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import oracle as O
    from taichimps_amd.workloads import ISOTROPIC as C
    w, theta, steps, p_dev, g_c, g_r, builds = block.pop("_check_inputs")
    u = np.random.default_rng([1004, 2]).normal(size=len(w["x"]))
    t0 = time.perf_counter()
    t, d, tl_builds = O.tlm_tangent(w, theta, steps, dradius=u, style="hertz/history",
                                    skin=C["skin"], dt=C["dt"], p_target=C["p_target"],
                                    p_gain=C["p_gain"])
    elapsed = time.perf_counter() - t0
    mean_p = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0]) / 3.0
    tl = mean_p @ d
    names = ("kn", "kt", "gamman", "gammat", "xmu")
    comp = {}
    for k, name in enumerate(names):
        comp[name] = {"adjoint": float(g_c[k]), "tangent_linear": float(tl[k]),
                      "rel_err": abs(g_c[k] - tl[k]) / abs(tl[k])}
    comp["radius_direction"] = {"adjoint": float(u @ g_r), "tangent_linear": float(tl[5]),
                                "rel_err": abs(u @ g_r - tl[5]) / abs(tl[5])}
    block["tlm_check"] = {
        "method": "tangent-linear (forward-mode) run of the same 2000 steps on the CPU "
                  "(oracle/tlm.cpp), FP64 values bitwise the device forward's",
        "forward_values_bitwise_equal": bool(float(mean_p @ t) == p_dev and tl_builds == builds),
        "dP": comp,
        "max_rel_err": max(c["rel_err"] for c in comp.values()),
        "threads": int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0))),
        "elapsed_s": elapsed,
    }


def build_member(w, precision):
    from taichimps_amd import workloads
    return workloads.build_member(w, precision)


def ensemble_c5(steps=500, n_grad=256, every=100):
    """Ensemble of 256 packings (taichimps_amd.workloads.ensemble: N log-uniform
    in [2048, 65536], per-member kn, xmu, target P, hooke/hertz alternating),
    FP32-split, `steps` steps after each member's setup force. Times the
    ensemble run (its autotuner's probe segments included) and the same
    members run one Simulation at a time, checks that every member's final
    state is bitwise the same both ways, and computes per-member adjoint
    gradients of the final mean pressure for `n_grad` members spread over the
    size range, all of them by default (FP64 twins: the adjoint engine
    differentiates the FP64 step)."""
    import torch
    from taichimps_amd import adjoint as ADJ
    from taichimps_amd import workloads
    from taichimps_amd.ensemble import Ensemble
    from taichimps_amd.workloads import ISOTROPIC as C
    ws = workloads.ensemble()
    natoms = int(sum(len(w["x"]) for w in ws))

    def setup(sims):
        for s in sims:
            s.run(0)
            s.run_pre = False

    sims = [build_member(w, "f32split") for w in ws]
    ens = Ensemble(sims)
    ens.run(0)
    for s in sims:
        s.run_pre = False
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run(steps, autotune=True)
    torch.cuda.synchronize()
    t_ens = time.perf_counter() - t0
    tuning = ens.tuning()
    print(f"# ensemble: {natoms} atoms, {steps} steps in {t_ens:.3f} s, {tuning}",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ens.run(steps)
    torch.cuda.synchronize()
    t_ens2 = time.perf_counter() - t0
    tuning2 = ens.tuning()

    solo = [build_member(w, "f32split") for w in ws]
    setup(solo)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in solo:
        s.run(steps)
    torch.cuda.synchronize()
    t_solo = time.perf_counter() - t0
    print(f"# one at a time: {t_solo:.3f} s", file=sys.stderr, flush=True)
    for s in solo:
        s.run(steps)
    differ = 0
    for a, b in zip(sims, solo):
        xa, xb = a.atom.x.to_numpy(), b.atom.x.to_numpy()
        va, vb = a.atom.v.to_numpy(), b.atom.v.to_numpy()
        if not (np.array_equal(xa, xb) and np.array_equal(va, vb)
                and np.array_equal(a.domain.boxhi, b.domain.boxhi)):
            differ += 1
    rebuilds = int(sum(s.neighbor.build_count for s in sims))
    del solo

    order = np.argsort([len(w["x"]) for w in ws])
    pick = order[np.linspace(0, len(ws) - 1, n_grad).round().astype(int)]
    mean_p = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0]) / 3.0
    from taichimps_amd.ensemble import gradients as member_gradients
    builders, thetas, radii = [], [], []
    for k in pick:
        w = ws[k]
        builders.append(ADJ.isotropic_builder(w, style=w["style"], skin=C["skin"], dt=C["dt"],
                                              p_target=w["p_target"], p_gain=C["p_gain"]))
        thetas.append(np.array([w["kn"], 2.0 / 7.0 * w["kn"], C["gamman"], C["gammat"],
                                w["xmu"]]))
        radii.append(w["radius"])
    print(f"# gradients of {len(pick)} members on 16 host threads", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    res = member_gradients(builders, thetas, radii, steps, every, mean_p, workers=16)
    t_grad = time.perf_counter() - t0
    grads = []
    for k, (tensor, gc, gr) in zip(pick, res):
        w = ws[k]
        grads.append({"member": int(k), "atoms": len(w["x"]), "style": w["style"],
                      "P": float(mean_p @ tensor),
                      "dP_dkn": float(gc[0]), "dP_dxmu": float(gc[4]),
                      "dP_dgamman": float(gc[2]),
                      "dP_dradius_norm": float(np.linalg.norm(gr))})
    print(f"# gradients of {len(pick)} members: {t_grad:.1f} s", file=sys.stderr, flush=True)
    return {
        "workload": f"{len(ws)} periodic packings, {natoms} atoms (N log-uniform in "
                    "[2048, 65536]), hooke/hertz history alternating, per-member kn, xmu, "
                    f"target P; FP32-split, {steps} steps",
        "atoms": natoms, "members": len(ws), "steps": steps,
        "atom_steps_per_s": natoms * steps / t_ens,
        "ms_per_step": t_ens / steps * 1e3,
        "tuned_rerun_atom_steps_per_s": natoms * steps / t_ens2,
        "one_at_a_time_atom_steps_per_s": natoms * steps / t_solo,
        "speedup_vs_one_at_a_time": t_solo / t_ens,
        "members_bitwise_equal_single_runs": len(ws) - differ,
        "tuning": tuning, "rerun_redone_steps": tuning2["redone_steps"],
        "rerun_redo_passes": tuning2["redo_passes"], "rerun_redo_ms": tuning2["redo_ms"],
        "list_builds": rebuilds,
        "gradients": {"count": len(grads), "steps": steps, "checkpoint_every": every,
                      "seconds": t_grad,
                      "by_size": [grads[q] for q in
                                  np.linspace(0, len(grads) - 1, min(8, len(grads)))
                                  .round().astype(int)],
                      "all_finite": bool(all(np.isfinite([g["dP_dkn"], g["dP_dxmu"],
                                                          g["dP_dgamman"]]).all()
                                             for g in grads)),
                      "workers": 16,
                      "note": "adjoint of each member's FP64 twin (same inputs, FP64 step), "
                              "members driven by 16 host threads at once"},
    }


def next_rows(dims=None, reps=20, warm_steps=20):
    """The widened rows on the 1M-sphere workload (FP64 system after
    `warm_steps` steps): per-contact computes, a floor wall's contact kernel,
    restart files, and the script driver on the 8,000-sphere case. API wall
    times include each call's device-to-host copy and synchronisation; kernel
    durations come from rocprofv3 (profiles/)."""
    import tempfile

    import taichimps_amd as tm
    from taichimps_amd import restart as RS
    from taichimps_amd import workloads
    from taichimps_amd.input import LAMMPSInputParser
    w = workloads.case("isotropic_hertz_1m", dims=tuple(dims) if dims else None)
    n = len(w["x"])
    sim = build_sim(w, "f64")
    sim.run(warm_steps)
    atom, dom, nl = sim.atom, sim.domain, sim.neighbor

    def timed(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3

    ca, fab, sa = tm.ComputeContactAtom(n), tm.ComputeFabric(), tm.ComputeStressAtom(n)
    out = {"workload": f"{n} spheres, isotropic compression (gran/hertz/history, FP64) after "
                       f"{warm_steps} steps", "atoms": n, "reps": reps}
    cnt = ca.compute(atom, dom, nl)
    phi, nc = fab.compute(atom, dom, nl)
    out["contacts"] = nc
    out["mean_coordination"] = float(cnt.mean())
    out["fabric_trace"] = float(phi[:3].sum())
    out["contact_atom_call_ms"] = timed(lambda: ca.compute(atom, dom, nl))
    out["fabric_call_ms"] = timed(lambda: fab.compute(atom, dom, nl))
    out["stress_atom_call_ms"] = timed(lambda: sa.compute(atom, dom, kinetic=True))
    out["stress_macro_call_ms"] = timed(lambda: sa.compute_macro(atom, dom, kinetic=True))
    # a floor under the lowest layer of spheres (centres 0.55 above the box's
    # lower face): about a tenth of the atoms touch it
    dom.pull()
    floor = float(dom.boxlo[2]) + 0.08
    wall = tm.FixWallGran(dom, 2, -1, floor, 2.0e5, 50.0, 2.0e5 * 2.0 / 7.0, 25.0, 0.5,
                          max_atoms=n)

    def wall_call():
        wall.post_force(atom, 1e-4)
        atom._native.sync()
    out["wall_post_force_call_ms"] = timed(wall_call)
    out["wall_touching_atoms"] = int(wall.touch.to_numpy()[:n].sum())
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "bench.restart")
        t0 = time.perf_counter()
        RS.write_restart(path, atom, dom, timestep=sim.timestep, dt=sim.dt,
                         history=sim.history, nlist=nl, pair_style="gran/hertz/history")
        out["write_restart_s"] = time.perf_counter() - t0
        out["restart_bytes"] = os.path.getsize(path)
        t0 = time.perf_counter()
        data = RS.read_restart(path)
        out["read_restart_s"] = time.perf_counter() - t0
        out["restart_atoms_read"] = len(data.tag)
        # script driver: the 8,000-sphere case as a script vs the same run via the API
        w8 = workloads.case("isotropic_hooke_8k")
        lines = ["LAMMPS data", "", f"{len(w8['x'])} atoms", "1 atom types", ""]
        for k, a in enumerate("xyz"):
            lines.append(f"{float(w8['boxlo'][k])!r} {float(w8['boxhi'][k])!r} {a}lo {a}hi")
        lines += ["", "Atoms # sphere", ""]
        for i in range(len(w8["x"])):
            x = w8["x"][i]
            lines.append(f"{i + 1} 1 {float(2 * w8['radius'][i])!r} 1.0 {float(x[0])!r} "
                         f"{float(x[1])!r} {float(x[2])!r}")
        lines += ["", "Velocities", ""]
        for i in range(len(w8["x"])):
            v = w8["v"][i]
            lines.append(f"{i + 1} {float(v[0])!r} {float(v[1])!r} {float(v[2])!r} 0 0 0")
        with open(os.path.join(tmp, "c1.data"), "w") as f:
            f.write("\n".join(lines) + "\n")
        with open(os.path.join(tmp, "in.c1"), "w") as f:
            f.write("atom_style sphere\nboundary p p p\nread_data c1.data\n"
                    "pair_style gran/hooke/history 2e5 NULL 50 NULL 0.5 1\npair_coeff * *\n"
                    "neighbor 0.1 bin\nneigh_modify delay 0 every 1 check yes\n"
                    "timestep 1e-4\nfix 1 all nve/sphere\nfix 2 all deform/pressure 1 "
                    "x pressure 1.0 1e-4 y pressure 1.0 1e-4 z pressure 1.0 1e-4 couple xyz\n"
                    "run 1000\n")
        t0 = time.perf_counter()
        LAMMPSInputParser(os.path.join(tmp, "in.c1")).execute()
        out["script_c1_1000_steps_s"] = time.perf_counter() - t0
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--prof-steps", type=int, default=50)
    ap.add_argument("--repeats", type=int, default=5,
                    help="fresh systems timed per precision; the median run is reported")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c2-200", action="store_true",
                    help="skip the 200-steps-from-the-start block")
    ap.add_argument("--no-walls", action="store_true",
                    help="skip the 1M-sphere run with walls inside the fused step")
    ap.add_argument("--walls-only", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint calibration block (32k spheres, 2000 steps)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only run the adjoint calibration block and print it")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the ensemble block (256 packings, 500 steps)")
    ap.add_argument("--triaxial", action="store_true",
                    help="run the 4M-sphere triaxial block (the default; kept for old commands)")
    ap.add_argument("--no-triaxial", action="store_true",
                    help="skip the 4M-sphere triaxial block (FP32-split and FP64, 1000 steps)")
    ap.add_argument("--triaxial-only", action="store_true",
                    help="only run the triaxial block and print it")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only run the ensemble block and print it")
    ap.add_argument("--ensemble-grads", type=int, default=256, metavar="N",
                    help="members whose adjoint gradients the ensemble block computes")
    ap.add_argument("--dims", type=int, nargs=3, default=None,
                    help="lattice dimensions (default 100 100 100)")
    ap.add_argument("--precision", choices=("both", "f64", "f32split"), default="both",
                    help="run one precision only (for profiler passes); the JSON line needs both")
    ap.add_argument("--direct", action="store_true",
                    help="launch each step kernel directly instead of replaying captured graphs")
    ap.add_argument("--next-rows-only", action="store_true",
                    help="only time the widened rows (contact computes, floor wall, restart "
                         "files, script driver) on the 1M-sphere workload and print them")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (a known-size "
                         "kernel that scales the profiler's HBM byte counters), then exit")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.adjoint_only:
        block = adjoint_c4()
        cpu_adjoint_check(block)
        print(json.dumps(block))
        return 0
    if args.ensemble_only:
        print(json.dumps(ensemble_c5(n_grad=args.ensemble_grads)))
        return 0
    if args.triaxial_only:
        print(json.dumps(triaxial_c3()))
        return 0
    if args.walls_only:
        print(json.dumps(wall_run()))
        return 0
    if args.next_rows_only:
        print(json.dumps(next_rows(dims=args.dims)))
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is a design only, not implemented)"}))
        return 2

    from taichimps_amd import _native, workloads
    w = workloads.case("isotropic_hertz_1m", dims=tuple(args.dims) if args.dims else None)
    device = _native.device_name()
    res = {}
    precisions = ("f64", "f32split") if args.precision == "both" else (args.precision,)
    for prec in precisions:
        res[prec] = run_precision(w, prec, args.steps, args.warmup, args.prof_steps, args.direct,
                                  repeats=args.repeats)
        r = res[prec]
        print(f"# {prec}: {r['ms_per_step']:.4f} ms/step, {r['atom_steps_per_s']:.4e} atom-steps/s, "
              f"kernels {r['kernel_ms']}, n_slot {r['n_slot']:.3f}, n_touch {r['n_touch']:.3f}, "
              f"rebuilds {r['rebuilds']}, P {r['mean_pressure']:.10g}", file=sys.stderr, flush=True)
    if args.precision != "both":
        return 0

    head = res["f32split"]
    f64 = res["f64"]
    n = head["n"]
    step_bytes = model_bytes(STEP_MODEL["f32split"], head["n_slot"], head["n_touch"]) * n
    kernel_ms = head["kernel_ms"]["step"]
    achieved = step_bytes / (kernel_ms * 1e-3) / 1e9
    step_achieved = step_bytes / (head["ms_per_step"] * 1e-3) / 1e9
    f64_step_bytes = model_bytes(STEP_MODEL["f64"], f64["n_slot"], f64["n_touch"]) * n

    c2 = None if args.no_c2_200 else c2_200()

    cpu = None
    if not args.no_cpu_baseline:
        # the serial oracle also checks the device's FP64 positions bit for bit
        cb = cpu_baseline(w, args.warmup, args.steps)
        ct = cpu_baseline(w, args.warmup, args.steps, threaded=True)
        w1 = workloads.case("isotropic_hooke_8k")
        c1t = cpu_baseline(w1, 10, 1000, threaded=True, style="hooke/history")
        c1s = cpu_baseline(w1, 10, 1000, style="hooke/history")
        cpu = {"value": ct["atom_steps_per_s"], "unit": "atom-steps/s", "cores": ct["threads"],
               "kind": "port",
               "sample": f"CPU oracle (C++ FP64 restatement; OpenMP build, {ct['threads']} "
                         f"threads, pair tally with atomics as the reference's kernels) on the "
                         f"same {n}-atom workload: {args.warmup} setup steps untimed, then "
                         f"{args.steps} steps timed ({ct['elapsed_s']:.1f} s)",
               "mean_pressure": ct["mean_pressure"],
               "single_thread": {"value": cb["atom_steps_per_s"], "cores": 1,
                                 "elapsed_s": cb["elapsed_s"]},
               "c1_isotropic_hooke_8k": {
                   "value": c1t["atom_steps_per_s"], "cores": c1t["threads"],
                   "single_thread": c1s["atom_steps_per_s"],
                   "sample": "8,000 spheres, gran/hooke/history, deform/pressure P=1.0: 10 setup "
                             "steps untimed, then 1,000 steps timed"},
               "host_cpus": len(os.sched_getaffinity(0))}
        if c2 is not None:
            oc = cpu_c2_200()
            for prec in ("f32split", "f64"):
                c2[prec]["mean_pressure_rel_err_vs_cpu_oracle"] = (
                    abs(c2[prec]["mean_pressure"] - oc["mean_pressure"]) / abs(oc["mean_pressure"]))
            c2["cpu_oracle_mean_pressure"] = oc["mean_pressure"]
            c2["cpu_oracle_rebuilds"] = oc["build_count"] - 1
            c2["fp64_positions_bitwise_equal_cpu_oracle"] = bool(
                np.array_equal(c2["f64_x"], oc["x"]))
    if c2 is not None:
        c2.pop("f64_x", None)

    out = {
        "metric": "atom-steps/s @1M Hertz/history: FP32-split vs FP64 plain; % HBM peak",
        "value": head["atom_steps_per_s"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "repeats": args.repeats,
        "ms_per_step_runs": head["ms_per_step_runs"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 contacts/integration, f64 bin origin + virial + servo (FP32-split)",
        "data": "synthetic (jittered 100^3 lattice, seed 1002)",
        "config": {"workload": "isotropic compression, 1M spheres, gran/hertz/history, "
                               "deform/pressure P=1.0, dt 1e-4, skin 0.1",
                   "atoms": n, "precision": "f32split", "device": device},
        "fp64": {"value": f64["atom_steps_per_s"], "ms_per_step": f64["ms_per_step"],
                 "kernel_ms": f64["kernel_ms"], "mean_pressure": f64["mean_pressure"],
                 "step_model_GBs": f64_step_bytes / (f64["ms_per_step"] * 1e-3) / 1e9,
                 "kernel_achieved_GBs": f64_step_bytes / (f64["kernel_ms"]["step"] * 1e-3) / 1e9},
        "speedup_fp32split_vs_fp64": head["atom_steps_per_s"] / f64["atom_steps_per_s"],
        "mean_pressure_rel_err_vs_fp64": abs(head["mean_pressure"] - f64["mean_pressure"])
        / abs(f64["mean_pressure"]),
        "roofline": {"bound": "hbm", "kernel": "k_step_split (one kernel per step)",
                     "limiter": "vector ALU issue and dependent neighbour-gather latency, not "
                                "HBM bandwidth (SQ and TCC counters in profiles/)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(),
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                     "copy-calibrated; profiles/)",
                     "bytes_per_launch": step_bytes, "avg_ms": kernel_ms},
        "step_roofline": {"achieved": step_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_achieved / HBM_PEAK_GBS,
                          "bytes_per_atom_step": step_bytes / n,
                          "note": "whole run incl. rebuilds and segment overhead"},
        "n_slot": head["n_slot"], "n_touch": head["n_touch"], "rebuilds": head["rebuilds"],
        "kernel_ms": head["kernel_ms"],
        "cpu_baseline": cpu,
        "c2_200": c2,
    }
    if not args.no_walls:
        out["walls"] = wall_run()
    if not args.no_adjoint:
        out["adjoint"] = adjoint_c4()
        if args.no_cpu_baseline:
            out["adjoint"].pop("_check_inputs")
        else:
            cpu_adjoint_check(out["adjoint"])
    if not args.no_ensemble:
        out["ensemble"] = ensemble_c5(n_grad=args.ensemble_grads)
    if not args.no_triaxial:
        out["triaxial"] = triaxial_c3()
    if cpu is not None:
        out["mean_pressure_rel_err_vs_cpu_oracle"] = (
            abs(head["mean_pressure"] - cpu["mean_pressure"]) / abs(cpu["mean_pressure"]))
        out["fp64_mean_pressure_rel_err_vs_cpu_oracle"] = (
            abs(f64["mean_pressure"] - cpu["mean_pressure"]) / abs(cpu["mean_pressure"]))
        out["fp64_positions_bitwise_equal_cpu_oracle"] = bool(np.array_equal(f64["x"], cb["x"]))
        out["f32split_max_abs_position_err_vs_cpu_oracle"] = float(np.abs(head["x"] - cb["x"]).max())
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
