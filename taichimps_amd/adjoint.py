"""Gradients of a DEM run's final stress with respect to the contact
coefficients and the per-particle radii.

No reference counterpart: the reference computes no gradients. The forward
is the FP64 device path with Simulation.run semantics (fix setup,
Verlet::setup, `nsteps` steps), checkpointed every `every` steps. The backward
is the hand-written adjoint of contact + velocity Verlet + the deform/pressure
servo in libtmps_hip.so (tmps_adjoint_forward / tmps_adjoint_backward).

    P, Z = final_pressure(coeffs, radius, builder, nsteps, every)

coeffs is a float64 tensor [kn, kt, gamman, gammat, xmu]. radius is a float64
tensor of the n radii (masses follow at the system's density). builder(coeffs,
radius) is a callable returning a fresh Simulation that uses them.
`P` is the final pressure tensor (xx yy zz xy xz yz, virial plus kinetic over
volume, as the servo saw it at the last step) and is differentiable.
`Z` is the final coordination number (touching contacts per atom). Z is
piecewise constant in every input, so its gradient is zero almost everywhere;
it is returned as a non-differentiable output.

Supported: gran/hooke/history and gran/hertz/history, fix nve/sphere, and fix
deform/pressure with the pressure and pressure/mean styles, remap x,
nevery 1. Anything else raises.
"""

from __future__ import annotations

from collections.abc import Callable

import numpy as np

COEFFS = ("kn", "kt", "gamman", "gammat", "xmu")


def run_forward(sim, nsteps: int, every: int):
    """Simulation.run(nsteps) on the device in FP64 with adjoint checkpoints
    every `every` steps; returns (final pressure tensor [6], coordination)."""
    from .EXTRA_FIX.deform_pressure import FixDeformPressure

    plan = sim._device_plan()
    if plan is None:
        raise NotImplementedError("the adjoint needs fix nve/sphere, a pair style, a neighbor "
                                  "list and contact history, with device-resident fixes only")
    cundall, deform, walls = plan
    if cundall is not None:
        raise NotImplementedError("the adjoint does not cover fix damping/cundall")
    if walls:
        raise NotImplementedError("the adjoint does not cover wall fixes")
    if not isinstance(deform, FixDeformPressure):
        raise NotImplementedError("the adjoint differentiates the pressure of fix deform/pressure")
    sim.setup_run(nsteps)
    atom = sim.atom
    N = atom._native
    sim.neighbor._bind(atom)
    sim.neighbor._push()
    sim.history._bind(atom)
    sim.pair_style._install(atom)
    N.set_counters(timestep=sim.timestep, setup_done=sim.setup_done)
    try:
        tensor, coordination = N.adjoint_forward(nsteps, every, sim.dt)
    finally:
        _sync_counters(sim)
    sim.domain.pull()
    return tensor, coordination


def run_backward(sim, dL_dtensor, initial_state: bool = False):
    """The adjoint of the last run_forward: dL/d(coeffs) [5], dL/d(radius) [n]
    and, when asked, dL/d(initial x, v, omega)."""
    n = sim.atom.nlocal
    try:
        return sim.atom._native.adjoint_backward(dL_dtensor, n, initial_state)
    finally:
        _sync_counters(sim)


def _sync_counters(sim):
    c = sim.atom._native.counters()
    sim.timestep = c["timestep"]
    sim.setup_done = bool(c["setup_done"])
    sim.neighbor._pull()
    for fix in sim.fixes:
        fix.timestep = sim.timestep
    if sim.integrator is not None:
        sim.integrator.timestep = sim.timestep
    sim.domain.mark_device_authoritative(sim.atom)


def _function():
    import torch

    class FinalPressure(torch.autograd.Function):
        """(coeffs, radius) -> (final pressure tensor, coordination number)."""

        @staticmethod
        def forward(ctx, coeffs, radius, builder, nsteps, every):
            c = coeffs.detach().to("cpu", torch.float64).numpy()
            r = radius.detach().to("cpu", torch.float64).numpy()
            sim = builder(c, r)
            tensor, z = run_forward(sim, int(nsteps), int(every))
            ctx.sim = sim
            ctx.device = coeffs.device
            P = torch.tensor(tensor, dtype=torch.float64, device=coeffs.device)
            Z = torch.tensor(z, dtype=torch.float64, device=coeffs.device)
            ctx.mark_non_differentiable(Z)
            return P, Z

        @staticmethod
        def backward(ctx, gP, gZ):
            g = gP.detach().to("cpu", torch.float64).numpy()
            gc, gr, _ = run_backward(ctx.sim, g)
            dev = ctx.device
            return (torch.from_numpy(gc).to(dev), torch.from_numpy(gr).to(dev), None, None, None)

    return FinalPressure


_FN = None


def final_pressure(coeffs, radius, builder: Callable, nsteps: int, every: int = 100):
    """Differentiable final pressure tensor (and coordination number) of the
    run `builder(coeffs, radius)` builds; see the module docstring."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(coeffs, radius, builder, nsteps, every)


def isotropic_builder(workload: dict, style: str = "hertz/history", skin: float = 0.1,
                      dt: float = 1e-4, p_target: float = 1.0, p_gain: float = 1e-4,
                      max_nb: int = 64) -> Callable:
    """A builder for the isotropic-compression runs: the workload's atoms
    (positions, velocities, spins, density, tags) with the given radii, the
    given legacy history pair style, nve/sphere and deform/pressure P."""
    import taichimps_amd as tm

    def build(coeffs, radius):
        kn, kt, gamman, gammat, xmu = (float(v) for v in coeffs)
        w = workload
        n = len(w["x"])
        dom = tm.Domain(w["boxlo"], w["boxhi"], ("p", "p", "p"))
        atom = tm.AtomSystem(max_atoms=n)
        atom.add_particles(x=w["x"], radius=np.asarray(radius, dtype=np.float64),
                           density=w["density"], v=w["v"], omega=w["omega"], tag=w["tag"])
        nl = tm.NeighborList(dom, skin=skin, max_atoms=n, max_neighbors_per_atom=max_nb)
        hist = tm.ContactHistory(n, max_nb)
        if style == "hooke/history":
            pair = tm.GranHookeHistory(dom, kn, gamman, kt, gammat, xmu)
        elif style == "hertz/history":
            pair = tm.GranHertzHistory(dom, kn, gamman, kt, gammat, xmu)
        else:
            raise NotImplementedError(f"the adjoint covers the legacy history styles, not {style}")
        sim = tm.Simulation(dom, atom, nl, hist, pair, dt=dt)
        sim.add_fix(tm.FixNVESphere(dom))
        sim.add_fix(tm.FixDeformPressure(dom, p_target=p_target, p_gain=p_gain))
        return sim

    return build
