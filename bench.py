"""Benchmark of the DEM step on one MI355X.

Workload: isotropic compression of 1,000,000 spheres (100^3 jittered lattice,
a = 1.1, radii U[0.45, 0.55], seed 1002), gran/hertz/history kn = 2e5,
kt = 2/7 kn, gamman = 50, gammat = 25, xmu = 0.5, dt = 1e-4, skin 0.1,
nve/sphere + deform/pressure isotropic P = 1.0 (pgain 1e-4, couple xyz,
remap x). Synthetic inputs from taichimps_amd.workloads.

For each precision (FP64 plain, FP32-split) the script runs `--warmup` steps
(the run's setup force and first neighbor build included), then times exactly
`--steps` device-resident steps (Simulation.run with `pre no`) between two
stream synchronisations. `value` is the FP32-split throughput in atom-steps/s;
the FP64 figure and their ratio ride along. A second pass of the same steps
with per-kernel HIP events gives the force kernel's average duration for the
roofline; the CPU baseline times the single-threaded CPU oracle on the same
workload.

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
    import taichimps_amd as tm
    from taichimps_amd.workloads import ISOTROPIC as C
    n = len(w["x"])
    dom = tm.Domain(w["boxlo"], w["boxhi"], ("p", "p", "p"))
    atom = tm.AtomSystem(max_atoms=n, float_type=precision)
    atom.add_particles(x=w["x"], radius=w["radius"], density=w["density"], v=w["v"],
                       omega=w["omega"], tag=w["tag"])
    nl = tm.NeighborList(dom, skin=C["skin"], max_atoms=n, max_neighbors_per_atom=64)
    hist = tm.ContactHistory(n, 64)
    pair = tm.GranHertzHistory(dom, C["kn"], C["gamman"], C["kt"], C["gammat"], C["xmu"])
    sim = tm.Simulation(dom, atom, nl, hist, pair, dt=C["dt"])
    sim.add_fix(tm.FixNVESphere(dom))
    sim.add_fix(tm.FixDeformPressure(dom, p_target=C["p_target"], p_gain=C["p_gain"]))
    return sim


# Algorithmic bytes per atom of each fused kernel (own-atom streams, list slots
# and touching-slot history; neighbour gathers are cache traffic and not
# counted): B = base + per_slot * n_slot + per_touch * n_touch.
BYTES = {
    # force pass: x|r r, v r+w, omega r+w, m, list offset (FP64: 24+8 + 48 + 48 + 8 + 4)
    ("f64", "force"): (140.0, 5.0, 48.0),
    # begin pass: x r+w, v r, x0 r (FP64: 48 + 24 + 24)
    ("f64", "begin"): (96.0, 0.0, 0.0),
    # force pass: u|r r, cell r, v r+w, omega r+w, m, list offset (12+4 + 4 + 24 + 24 + 4 + 4)
    ("f32split", "force"): (76.0, 5.0, 24.0),
    # begin pass: u r+w, cell r, v r, u0 r (24 + 4 + 12 + 12)
    ("f32split", "begin"): (52.0, 0.0, 0.0),
}
# Whole-step model (per atom-step): B = B_atom + 6 n_slot + B_hist n_touch
STEP_MODEL = {"f64": (188.0, 6.0, 48.0), "f32split": (100.0, 6.0, 24.0)}


def model_bytes(key, n_slot, n_touch):
    base, per_slot, per_touch = key
    return base + per_slot * n_slot + per_touch * n_touch


def run_precision(w, precision, steps, warmup, prof_steps):
    sim = build_sim(w, precision)
    N = sim.atom._native
    n = sim.atom.nlocal
    t0 = time.perf_counter()
    sim.run(warmup)
    N.sync()
    t_setup = time.perf_counter() - t0
    sim.run_pre = False
    builds0 = sim.neighbor.build_count
    N.sync()
    t0 = time.perf_counter()
    sim.run(steps)
    N.sync()
    elapsed = time.perf_counter() - t0
    c = N.counters()
    rebuilds = sim.neighbor.build_count - builds0
    pressure = N.last_pressure()
    x_final = sim.atom.x.to_numpy()[:n]
    n_slot = c["nslots"] / n
    n_touch = c["ntouch"] / n
    # per-kernel HIP-event timings over the same number of steps, same stream
    N.set_profiling(True)
    sim.run(prof_steps)
    ms, cnt = N.kernel_times()
    N.set_profiling(False)
    return {
        "n": n, "elapsed_s": elapsed, "ms_per_step": elapsed / steps * 1e3,
        "atom_steps_per_s": n * steps / elapsed, "setup_s": t_setup, "rebuilds": rebuilds,
        "n_slot": n_slot, "n_touch": n_touch, "pressure": pressure.tolist(),
        "mean_pressure": float(pressure[:3].mean()),
        "kernel_ms": {"begin": ms[0], "force": ms[1], "end": ms[2]},
        "kernel_counts": cnt.tolist(), "x": x_final,
    }


def cpu_baseline(w, warmup, steps):
    """The single-threaded CPU oracle on the same workload, with the same run
    split as the device (a `warmup` run with setup, then `steps` timed)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import oracle as O
    from taichimps_amd.workloads import ISOTROPIC as C
    O.build()
    o = O.from_workload(w)
    o.set_neigh(C["skin"])
    o.set_pair(O.legacy_pair("hertz/history", kn=C["kn"], gamman=C["gamman"], kt=C["kt"],
                             gammat=C["gammat"], xmu=C["xmu"]))
    o.set_servo(O.servo_params(p_target=C["p_target"], p_gain=C["p_gain"]))
    assert o.run(warmup, C["dt"]) == 0
    t0 = time.perf_counter()
    assert o.run(steps, C["dt"], run_pre=False) == 0
    elapsed = time.perf_counter() - t0
    p = o.last_pressure_tensor()
    return {"atom_steps_per_s": len(w["x"]) * steps / elapsed, "elapsed_s": elapsed,
            "mean_pressure": float(p[:3].mean()), "x": o.atoms()["x"]}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--prof-steps", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--dims", type=int, nargs=3, default=None,
                    help="lattice dimensions (default 100 100 100)")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is described in DESIGN.md, not implemented)"}))
        return 2

    from taichimps_amd import _native, workloads
    w = workloads.case("isotropic_hertz_1m", dims=tuple(args.dims) if args.dims else None)
    device = _native.device_name()
    res = {}
    for prec in ("f64", "f32split"):
        res[prec] = run_precision(w, prec, args.steps, args.warmup, args.prof_steps)
        r = res[prec]
        print(f"# {prec}: {r['ms_per_step']:.4f} ms/step, {r['atom_steps_per_s']:.4e} atom-steps/s, "
              f"kernels {r['kernel_ms']}, n_slot {r['n_slot']:.3f}, n_touch {r['n_touch']:.3f}, "
              f"rebuilds {r['rebuilds']}, P {r['mean_pressure']:.10g}", file=sys.stderr, flush=True)

    head = res["f32split"]
    f64 = res["f64"]
    n = head["n"]
    force_bytes = model_bytes(BYTES[("f32split", "force")], head["n_slot"], head["n_touch"]) * n
    force_ms = head["kernel_ms"]["force"]
    achieved = force_bytes / (force_ms * 1e-3) / 1e9
    step_bytes = model_bytes(STEP_MODEL["f32split"], head["n_slot"], head["n_touch"]) * n
    step_achieved = step_bytes / (head["ms_per_step"] * 1e-3) / 1e9
    f64_step_bytes = model_bytes(STEP_MODEL["f64"], f64["n_slot"], f64["n_touch"]) * n
    f64_force_bytes = model_bytes(BYTES[("f64", "force")], f64["n_slot"], f64["n_touch"]) * n

    cpu = None
    if not args.no_cpu_baseline:
        cb = cpu_baseline(w, args.warmup, args.steps)
        cpu = {"value": cb["atom_steps_per_s"], "unit": "atom-steps/s", "cores": 1, "kind": "port",
               "sample": f"CPU oracle (C++ FP64 restatement, 1 thread) on the same {n}-atom "
                         f"workload: {args.warmup} setup steps untimed, then {args.steps} steps "
                         f"timed ({cb['elapsed_s']:.1f} s)",
               "mean_pressure": cb["mean_pressure"]}

    out = {
        "metric": "atom-steps/s @1M Hertz/history: FP32-split vs FP64 plain; % HBM peak",
        "value": head["atom_steps_per_s"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
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
                 "force_achieved_GBs": f64_force_bytes / (f64["kernel_ms"]["force"] * 1e-3) / 1e9},
        "speedup_fp32split_vs_fp64": head["atom_steps_per_s"] / f64["atom_steps_per_s"],
        "mean_pressure_rel_err_vs_fp64": abs(head["mean_pressure"] - f64["mean_pressure"])
        / abs(f64["mean_pressure"]),
        "roofline": {"bound": "hbm", "kernel": "k_step_force_split (force pass)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": force_bytes, "avg_ms": force_ms},
        "step_roofline": {"achieved": step_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_achieved / HBM_PEAK_GBS,
                          "bytes_per_atom_step": step_bytes / n},
        "n_slot": head["n_slot"], "n_touch": head["n_touch"], "rebuilds": head["rebuilds"],
        "kernel_ms": head["kernel_ms"],
        "cpu_baseline": cpu,
    }
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
