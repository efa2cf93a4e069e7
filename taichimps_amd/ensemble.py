"""Many independent Simulations advanced together on the device.

No reference counterpart: the reference runs one Simulation at a time. An
Ensemble takes FP32-split Simulations whose runs the device path covers
(nve/sphere, a pair style with contact history, optionally deform/pressure;
no fix damping/cundall) and advances all of them with one launch per step and
pair style (libtmps_hip.so, tmps_batch_*); the steps that members redo with a
list rebuild are redone together, one launch per kernel of the redo for all
of them. Every member's state after ``Ensemble.run(steps)`` is bitwise the
state ``member.run(steps)`` gives it alone; members may differ in size, box,
pair style, coefficients and servo target, but share the time step.

    ens = Ensemble([sim_a, sim_b, ...])
    ens.run(500, autotune=True)
    ens.tuning()   # the launch configuration the autotuner chose

gradients(): per-member adjoint gradients of the final pressure, from the
FP64 adjoint engine (taichimps_amd.adjoint) on an FP64 twin of the member.
"""

from __future__ import annotations

from collections.abc import Callable, Sequence

import numpy as np

from . import _native
from .EXTRA_FIX.deform_pressure import FixDeformPressure

SEGMENT_LENGTHS = (8, 8, 16, 16, 32, 32)
BLOCK_ORDERS = ("sequence", "xcd", "sequence", "xcd", "sequence", "xcd")


class Ensemble:
    def __init__(self, simulations: Sequence):
        sims = list(simulations)
        if not sims:
            raise ValueError("an ensemble needs at least one Simulation")
        if len({id(s) for s in sims}) != len(sims):
            raise ValueError("a Simulation appears twice in the ensemble")
        for s in sims:
            if s.atom.precision != _native.FP32_SPLIT:
                raise NotImplementedError("an ensemble runs FP32-split Simulations")
            plan = s._device_plan()
            if plan is None:
                raise NotImplementedError("an ensemble member needs fix nve/sphere, a pair style, "
                                          "a neighbor list and contact history, with device "
                                          "fixes only")
            if plan[0] is not None:
                raise NotImplementedError("ensembles do not run fix damping/cundall")
            if plan[2]:
                raise NotImplementedError("ensembles do not run wall fixes")
            if s.atom.nlocal < 1:
                raise ValueError("every ensemble member needs atoms")
        dts = {float(s.dt) for s in sims}
        if len(dts) != 1:
            raise ValueError("ensemble members share one time step")
        self.simulations = sims
        self._batch = None
        self._counts = None

    def _prepare(self, sim, steps):
        """The device-run prologue of Simulation.run for one member."""
        sim.setup_run(steps)
        atom = sim.atom
        N = atom._native
        sim.neighbor._bind(atom)
        sim.neighbor._push()
        sim.history._bind(atom)
        sim.pair_style._install(atom)
        _, deform, _ = sim._device_plan()
        if deform is None:
            off = _native.ServoParams()
            off.nevery = 1
            N.servo_set(off)
        N.set_counters(timestep=sim.timestep, setup_done=sim.setup_done)
        if sim.run_pre or not sim.setup_done:
            # Verlet::setup for this member alone (tmps_run with no steps)
            N.run(0, sim.dt, run_pre=True)

    def _sync(self, sim):
        c = sim.atom._native.counters()
        sim.timestep = c["timestep"]
        sim.setup_done = bool(c["setup_done"])
        sim.neighbor._pull()
        for fix in sim.fixes:
            fix.timestep = sim.timestep
        if sim.integrator is not None:
            sim.integrator.timestep = sim.timestep
        if any(type(f) is FixDeformPressure for f in sim.fixes):
            sim.domain.mark_device_authoritative(sim.atom)
        sim.last_run_device_resident = True

    def run(self, steps: int, autotune: bool = False) -> None:
        """Simulation.run(steps) on every member."""
        if steps < 0:
            raise ValueError("negative step count")
        sims = self.simulations
        try:
            for s in sims:
                self._prepare(s, steps)
            counts = tuple(s.atom.nlocal for s in sims)
            if self._batch is None or counts != self._counts:
                if self._batch is not None:
                    self._batch.close()
                self._batch = _native.NativeBatch([s.atom._native for s in sims])
                self._counts = counts
            self._batch.run(steps, sims[0].dt, run_pre=False, autotune=autotune)
        finally:
            for s in sims:
                self._sync(s)
        for s in sims:
            s.domain.pull()

    def tuning(self) -> dict:
        """The launch configuration in use and what the autotuner measured."""
        if self._batch is None:
            raise RuntimeError("the ensemble has not run yet")
        seg, order, rates, redone, passes, redo_ms = self._batch.tuning()
        return {
            "segment_length": seg,
            "block_order": "xcd" if order else "sequence",
            "rates_atom_steps_per_s": {f"{SEGMENT_LENGTHS[k]}/{BLOCK_ORDERS[k]}": float(rates[k])
                                       for k in range(len(rates))},
            "redone_steps": redone,
            "redo_passes": passes,
            "redo_ms": redo_ms,
        }

    def close(self) -> None:
        if self._batch is not None:
            self._batch.close()
            self._batch = None


def gradients(builders: Sequence[Callable], coeffs: Sequence, radii: Sequence, nsteps: int,
              every: int, dL_dtensor=None, workers: int = 16) -> list:
    """Per-member gradients of the final pressure: for member k, builder k
    (coeffs k [kn, kt, gamman, gammat, xmu], radii k) builds its FP64 twin,
    whose run the adjoint engine reverses. dL_dtensor defaults to the mean
    pressure (1/3 on the diagonal). Returns a list of
    (final pressure tensor, dL/dcoeffs [5], dL/dradius [n]).

    Members are independent systems on streams of their own, so `workers`
    host threads drive their forward and backward passes at once (the
    library calls release the interpreter lock) and the GPU runs several
    members' kernels together; each member's result does not depend on it."""
    from concurrent.futures import ThreadPoolExecutor

    from . import adjoint as ADJ

    g = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0]) / 3.0 if dL_dtensor is None else dL_dtensor

    def one(item):
        build, c, r = item
        sim = build(np.asarray(c, dtype=np.float64), np.asarray(r, dtype=np.float64))
        tensor, _ = ADJ.run_forward(sim, nsteps, every)
        gc, gr, _ = ADJ.run_backward(sim, g)
        del sim
        return tensor, gc, gr

    items = list(zip(builders, coeffs, radii))
    if workers <= 1 or len(items) <= 1:
        return [one(it) for it in items]
    with ThreadPoolExecutor(max_workers=workers) as pool:
        return list(pool.map(one, items))
