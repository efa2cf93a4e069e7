"""fix print (reference: EXTRA_FIX/print.py:15-70, LAMMPS fix_print.cpp).

`fix ID group print N "text" [file path] [screen yes|no] [title "t"]`: on every
timestep that is a multiple of N (the absolute step, so a run resumed from a
restart prints on the same steps as an unbroken one), the text with its
variables substituted is written as one line to the file and/or the screen.

The fix reads the state and changes nothing, so a run that carries it stays
device-resident: Simulation cuts the device run at the multiples of N and
calls ``advance`` there, instead of the per-step ``end_of_step`` of the stage
path.
"""

from __future__ import annotations

from collections.abc import Callable
from pathlib import Path
from typing import Any

from .base import Fix


class FixPrint(Fix):
    """Periodic text output of evaluated variables (print.py:15-47)."""

    # reads only: a device-resident run is cut at its multiples, not given up
    output_only = True

    def __init__(self, domain, nevery: int, template_str: str, eval_fn: Callable[[str], str],
                 filepath: str | Path | None = None, title: str | None = None,
                 screen: bool = False, fix_id: str = "") -> None:
        super().__init__(domain)
        self.nevery = max(1, int(nevery))
        self.template_str = template_str
        self.eval_fn = eval_fn
        self.filepath = Path(filepath) if filepath else None
        self.title = title
        self.screen = screen
        self.file_handle: Any = None
        self.step_count = 0
        self.lines: list[str] = []
        if self.filepath:
            self.file_handle = open(self.filepath, "w")  # noqa: SIM115
            # a blank title writes no header line; no title writes the default
            if title is not None and title.strip() != "":
                self.file_handle.write(title.strip("\"'") + "\n")
            elif title is None:
                self.file_handle.write("# Fix print output" + (f" for fix {fix_id}" if fix_id
                                                               else "") + "\n")
            self.file_handle.flush()

    def advance(self, nsteps: int) -> None:
        """nsteps steps ended at self.timestep: print if it is a multiple of N
        (the device run is cut so no multiple falls inside the nsteps)."""
        self.step_count += int(nsteps)
        if self.timestep % self.nevery != 0:
            return
        text = self.eval_fn(self.template_str)
        self.lines.append(text)
        if self.file_handle:
            self.file_handle.write(text + "\n")
            self.file_handle.flush()
        if self.screen:
            print(text, flush=True)

    def end_of_step(self, atom, dt: float) -> None:
        """The stage path's per-step hook (print.py:49-66)."""
        self.advance(1)

    def close(self) -> None:
        if self.file_handle and not self.file_handle.closed:
            self.file_handle.close()

    def __del__(self) -> None:
        self.close()
