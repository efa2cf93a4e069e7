"""LAMMPS data files for atom_style sphere (reference: data_reader.py,
LAMMPS read_data.cpp).

Header lines give the atom count, the atom types and the box; the Atoms
section holds `id type diameter density x y z [ix iy iz]`, the optional
Velocities section `id vx vy vz [wx wy wz]`, matched to atoms by id. Other
sections are skipped.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np


@dataclass
class LAMMPSData:
    natoms: int
    natom_types: int
    boxlo: list[float]
    boxhi: list[float]
    tilt: list[float]
    x: np.ndarray
    radius: np.ndarray
    density: np.ndarray
    v: np.ndarray
    omega: np.ndarray
    tag: np.ndarray
    atom_type: np.ndarray


_SECTIONS = ("atoms", "velocities", "masses", "pair coeffs", "bonds", "angles")


def read_data(filepath: str | Path) -> LAMMPSData:
    """Parse a data file (FileNotFoundError when it does not exist)."""
    path = Path(filepath)
    if not path.exists():
        raise FileNotFoundError(f"Data file not found: {path}")
    section = "header"
    lines: dict[str, list[str]] = {"header": []}
    with open(path, encoding="utf-8", errors="ignore") as f:
        for raw in f:
            line = raw.split("#")[0].strip()
            if not line:
                continue
            low = line.lower()
            hit = next((s for s in _SECTIONS if low.startswith(s)), None)
            if hit is not None:
                section = hit
                lines[section] = []
                continue
            lines[section].append(line)

    natoms, ntypes = 0, 1
    lo, hi, tilt = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.0, 0.0, 0.0]
    for line in lines["header"]:
        t = line.split()
        if len(t) >= 2 and t[1] == "atoms":
            natoms = int(t[0])
        elif len(t) >= 3 and t[1:3] == ["atom", "types"]:
            ntypes = int(t[0])
        elif len(t) >= 4 and t[2][1:] == "lo" and t[3][1:] == "hi" and t[2][0] in "xyz":
            k = "xyz".index(t[2][0])
            lo[k], hi[k] = float(t[0]), float(t[1])
        elif len(t) >= 6 and t[3:6] == ["xy", "xz", "yz"]:
            tilt = [float(v) for v in t[0:3]]

    atoms = [ln.split() for ln in lines.get("atoms", [])]
    n = len(atoms)
    if natoms == 0:
        natoms = n
    if n:
        a = np.array([[float(v) for v in row[:7]] for row in atoms])
    else:
        a = np.zeros((0, 7))
    tag = a[:, 0].astype(np.int32)
    data = LAMMPSData(
        natoms=natoms, natom_types=ntypes, boxlo=lo, boxhi=hi, tilt=tilt,
        x=a[:, 4:7].copy(), radius=0.5 * a[:, 2], density=a[:, 3].copy(),
        v=np.zeros((n, 3)), omega=np.zeros((n, 3)), tag=tag,
        atom_type=a[:, 1].astype(np.int32))
    index = {int(t): i for i, t in enumerate(tag)}
    for ln in lines.get("velocities", []):
        t = ln.split()
        i = index.get(int(t[0]))
        if i is None:
            continue
        data.v[i] = [float(v) for v in t[1:4]]
        if len(t) >= 7:
            data.omega[i] = [float(v) for v in t[4:7]]
    return data
