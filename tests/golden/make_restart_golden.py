"""Writes the restart-file fixtures G6 with the reference's own writer.

    python tests/golden/make_restart_golden.py

The reference's restart module (src/taichimps/restart.py) depends on struct,
dataclasses, pathlib and numpy only, so it is loaded by file path from the
reference checkout (its package __init__ needs Taichi, which is absent). Its
write_restart is called on the inputs of restart_case.py:

g6_restart_full.bin   pair granular (hertz/material + mindlin +
                      coeff_restitution) and the contact history
g6_restart_plain.bin  no force field, no history

Only the bytes are kept: tests/test_restart.py checks that this package's
writer produces them byte for byte from the same inputs, and that its reader
recovers the inputs from them.
"""

from __future__ import annotations

import importlib.util
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import restart_case as R  # noqa: E402

REF = "/root/reference/src/taichimps/restart.py"


def main() -> None:
    spec = importlib.util.spec_from_file_location("reference_restart", REF)
    ref = importlib.util.module_from_spec(spec)
    sys.modules["reference_restart"] = ref  # dataclasses look their module up
    spec.loader.exec_module(ref)
    d = R.arrays()
    atom, dom, nl, hist = R.Atom(d), R.Domain(), R.NeighborList(d), R.History(d)
    with tempfile.TemporaryDirectory() as tmp:
        full = os.path.join(tmp, "full")
        ref.write_restart(full, atom, dom, timestep=R.TIMESTEP, dt=R.DT, history=hist,
                          pair=R.Pair(), pair_style="granular", nlist=nl)
        plain = os.path.join(tmp, "plain")
        ref.write_restart(plain, atom, dom, timestep=R.TIMESTEP, dt=R.DT)
        for src, name in ((full, "g6_restart_full.bin"), (plain, "g6_restart_plain.bin")):
            with open(src, "rb") as f:
                data = f.read()
            with open(os.path.join(HERE, name), "wb") as f:
                f.write(data)
            print(f"{name}: {len(data)} bytes")


if __name__ == "__main__":
    main()
