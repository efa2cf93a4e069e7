"""Simulation orchestrator (reference: simulation.py).

``run()`` executes LAMMPS Verlet ordering. When every fix is one the device
step implements (nve/sphere, damping/cundall, deform/pressure), the whole run
is device-resident (tmps_run: captured HIP-graph segments, rebuilds decided on
the device); otherwise each stage is launched from Python in the reference's
order (``step()``), each stage still a HIP kernel.
"""

from __future__ import annotations

from typing import Any

from . import _native
from .atom import AtomSystem
from .computes.thermo import Computes
from .constants import f64
from .contact_history import ContactHistory
from .domain import Domain
from .EXTRA_FIX.base import Fix
from .EXTRA_FIX.damping_cundall import FixDampingCundall
from .EXTRA_FIX.deform_pressure import FixDeformPressure
from .EXTRA_FIX.nve_sphere import FixNVESphere
from .EXTRA_FIX.wall_gran import FixWallGran
from .GRANULAR.base import GranularPair
from .neighbor import NeighborList


class Simulation:
    def __init__(
        self,
        domain: Domain,
        atom: AtomSystem | None = None,
        neighbor: NeighborList | None = None,
        history: ContactHistory | None = None,
        pair: GranularPair | None = None,
        max_atoms: int = 100000,
        max_neighbors_per_atom: int = 64,
        skin: float = 0.001,
        dt: float = 1e-4,
        thermo_freq: int = 100,
        float_type: Any = f64,
    ) -> None:
        self.domain = domain
        self.dt = dt
        self.float_type = float_type
        self.timestep = 0
        self.atom = atom if atom is not None else AtomSystem(max_atoms=max_atoms,
                                                             float_type=float_type)
        self.neighbor = neighbor if neighbor is not None else NeighborList(
            domain=domain, max_atoms=max_atoms, max_neighbors_per_atom=max_neighbors_per_atom,
            skin=skin, float_type=float_type)
        self.history = history if history is not None else ContactHistory(
            max_atoms=max_atoms, max_neighbors=max_neighbors_per_atom, float_type=float_type)
        self.pair_style: GranularPair | None = pair
        self.integrator: FixNVESphere | None = None
        self.fixes: list[Fix] = []
        self.setup_done = False
        self.run_pre = True
        # periodic restart output (`restart N file`, simulation.py:74-98)
        self.restart_every = 0
        self.restart_file: str | None = None
        self.restart_writer = None
        self.thermo_freq = thermo_freq
        self.computes = Computes(float_type=float_type)
        self.last_run_device_resident = False
        self.atom._bind_domain(domain)

    def set_pair_style(self, pair: GranularPair) -> None:
        self.pair_style = pair

    def add_fix(self, fix: Fix) -> None:
        if isinstance(fix, FixNVESphere):
            self.integrator = fix
        else:
            self.fixes.append(fix)

    def remove_fix(self, fix: Fix) -> None:
        if fix in self.fixes:
            self.fixes.remove(fix)
        elif self.integrator is fix:
            self.integrator = None

    # ------------------------------------------------------------ stages
    def reneighbor_if_needed(self, force: bool = False) -> bool:
        """save -> pbc -> build -> restore on LAMMPS's reneighbor steps (simulation.py:122-159)."""
        if self.neighbor is None:
            return False
        forced = force
        for fix in self.fixes:
            if fix.force_reneighbor:
                forced = True
                fix.force_reneighbor = False
        if not self.neighbor.decide(self.atom, self.timestep) and not forced:
            return False
        if self.history is not None:
            self.history.save_state(self.atom, self.neighbor)
        self.domain.pbc(self.atom)
        self.domain.pull()
        self.neighbor.build(self.atom)
        if self.history is not None:
            self.history.restore_state(self.atom, self.neighbor)
        return True

    def setup_run(self, nsteps_total: int) -> None:
        for fix in self.fixes:
            if hasattr(fix, "note_atom_count"):
                fix.note_atom_count(self.atom.nlocal)
            fix.setup(nsteps_total, self.dt)
        if self.integrator is not None:
            self.integrator.setup(nsteps_total, self.dt)

    def init_simulation(self) -> None:
        """Verlet::setup: rebuild and evaluate forces without advancing history (:170-203)."""
        self.setup_done = True
        self.atom.clear_forces()
        if self.pair_style is not None:
            self.reneighbor_if_needed(force=True)
            self.pair_style.compute(self.atom, self.neighbor, self.history, self.dt,
                                    shearupdate=False)
        for fix in self.fixes:
            fix.history_update = False
        for fix in self.fixes:
            fix.post_force(self.atom, self.dt)
        for fix in self.fixes:
            fix.history_update = True

    def _sub_step_inner(self) -> None:
        """One Verlet step in the reference's stage order (simulation.py:205-247)."""
        self.timestep += 1
        for fix in self.fixes:
            fix.timestep = self.timestep
        if self.integrator is not None:
            self.integrator.timestep = self.timestep
        if self.integrator is not None:
            self.integrator.initial_integrate(self.atom, self.dt)
        self.reneighbor_if_needed()
        self.atom.clear_forces()
        if self.pair_style is not None and self.neighbor is not None:
            self.pair_style.compute(self.atom, self.neighbor, self.history, self.dt)
        for fix in self.fixes:
            fix.post_force(self.atom, self.dt)
        if self.integrator is not None:
            self.integrator.final_integrate(self.atom, self.dt)
        for fix in self.fixes:
            fix.end_of_step(self.atom, self.dt)

    def step(self) -> None:
        self._sub_step_inner()

    # --------------------------------------------------------------- runs
    def _device_plan(self):
        """(cundall, deform, walls) when the run can stay on the device, else
        None. walls: up to two non-wiggling wall fixes, each with whether it
        comes before damping/cundall among the fixes (their post_force order,
        simulation.py:236-247)."""
        if not isinstance(self.integrator, FixNVESphere):
            return None
        if self.pair_style is None or self.neighbor is None or self.history is None:
            return None
        cundall = deform = None
        walls = []
        for fix in self.fixes:
            if getattr(fix, "output_only", False):
                continue  # fix print: the run is cut at its multiples
            if type(fix) is FixDampingCundall and cundall is None:
                cundall = fix
            elif type(fix) is FixDeformPressure and deform is None:
                deform = fix
            elif (isinstance(fix, FixWallGran) and fix.wiggle_amplitude == 0.0
                  and len(walls) < 2):
                walls.append((fix, cundall is None))
            else:
                return None
            if fix.force_reneighbor:
                return None
        return cundall, deform, walls

    def _set_walls(self, walls) -> None:
        atom = self.atom
        spec = []
        for fix, before in walls:
            w, coord, vs = fix._fused(atom)
            spec.append((w, before, coord, vs))
        atom._native.set_run_walls(spec, self.dt)

    def _run_device(self, steps: int, cundall, deform, walls) -> None:
        atom = self.atom
        N = atom._native
        self.neighbor._bind(atom)
        self.neighbor._push()
        self.history._bind(atom)
        self.history._push_seed(atom)
        builds_before = N.counters()["build_count"]
        self.pair_style._install(atom)
        if deform is None:
            off = _native.ServoParams()
            off.nevery = 1
            N.servo_set(off)
        N.set_counters(timestep=self.timestep, setup_done=self.setup_done)
        run_pre = self.run_pre or not self.setup_done
        gl_ga = None if cundall is None else (cundall.gamma_lin, cundall.gamma_ang)
        self._set_walls(walls)
        try:
            N.run(steps, self.dt, run_pre=run_pre, cundall=gl_ga)
            for fix, _ in walls:
                fix._advance(steps + (1 if run_pre else 0), self.dt)
        finally:
            N.set_run_walls([], self.dt)
            c = N.counters()
            self.timestep = c["timestep"]
            self.setup_done = bool(c["setup_done"])
            if c["build_count"] > builds_before:
                # the run's first build carried any restart seed onto its list
                self.history.setup_pending = False
            self.neighbor._pull()
            for fix in self.fixes:
                fix.timestep = self.timestep
            if self.integrator is not None:
                self.integrator.timestep = self.timestep
            if deform is not None:
                self.domain.mark_device_authoritative(atom)
        self.last_run_device_resident = True

    def set_restart(self, every: int, path: str | None, writer) -> None:
        """Periodic restart output, `restart N file` (simulation.py:86-90):
        writer(path, timestep) at every multiple of `every` during runs."""
        self.restart_every = int(every)
        self.restart_file = path
        self.restart_writer = writer

    def write_restart_now(self) -> None:
        """One periodic restart file for the current timestep (:92-98)."""
        if self.restart_writer is None or not self.restart_file:
            return
        self.restart_writer(self.restart_file.replace("*", str(self.timestep)), self.timestep)

    def _restart_due(self) -> bool:
        return bool(self.restart_every and self.restart_writer is not None
                    and self.timestep % self.restart_every == 0)

    def run(self, steps: int) -> None:
        """Run `steps` Verlet steps (simulation.py:320-332)."""
        self.setup_run(steps)
        plan = self._device_plan()
        if plan is not None and self.atom.nlocal > 0:
            if self._intervals():
                self._run_device_chunked(steps, plan)
            else:
                self._run_device(steps, *plan)
        else:
            self.last_run_device_resident = False
            if self.run_pre or not self.setup_done:
                self.init_simulation()
            for _ in range(steps):
                self.step()
                if self._restart_due():
                    self.domain.pull()
                    self.write_restart_now()
        self.domain.pull()

    def prepare_run(self, steps: int) -> None:
        """Capture the device graphs a following device-resident run(steps)
        launches, without running anything (no reference counterpart). A no-op
        before the first run or where the run takes the stage path."""
        plan = self._device_plan()
        if plan is None or self.atom.nlocal == 0 or not self.setup_done:
            return
        cundall, _, walls = plan
        gl_ga = None if cundall is None else (cundall.gamma_lin, cundall.gamma_ang)
        self._set_walls(walls)
        try:
            self.atom._native.prepare_run(steps, self.dt, cundall=gl_ga)
        finally:
            self.atom._native.set_run_walls([], self.dt)

    def _output_fixes(self) -> list:
        return [f for f in self.fixes if getattr(f, "output_only", False)]

    def _intervals(self) -> list[int]:
        """The I/O periods a device run is cut at: periodic restarts and
        output-only fixes (fix print's N), as run_gpu collects them
        (simulation.py:264-270)."""
        out = [f.nevery for f in self._output_fixes() if f.nevery > 0]
        if self.restart_every and self.restart_writer is not None:
            out.append(self.restart_every)
        return out

    def _run_device_chunked(self, steps: int, plan) -> None:
        """The device run cut at the I/O multiples (as run_gpu cuts it,
        simulation.py:249-292): one fix setup for the whole run, later pieces
        continuing without Verlet::setup, so the trajectory is the unbroken
        run's. At each cut the output fixes see the steps that passed (their
        end_of_step comes inside the step, before the restart write)."""
        saved = self.run_pre
        remaining = steps
        periods = self._intervals()
        try:
            while remaining > 0:
                chunk = min([remaining] + [p - self.timestep % p for p in periods])
                self._run_device(chunk, *plan)
                self.run_pre = False
                remaining -= chunk
                for fix in self._output_fixes():
                    fix.timestep = self.timestep
                    fix.advance(chunk)
                if self._restart_due():
                    self.domain.pull()
                    self.write_restart_now()
        finally:
            self.run_pre = saved

    def run_gpu(self, steps: int) -> None:
        """The reference's batched variant (simulation.py:249-292); identical here."""
        self.run(steps)
