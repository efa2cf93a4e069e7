"""LAMMPS binary restart files for atom_style sphere (reference: restart.py).

The on-disk layout is LAMMPS's own (write_restart.cpp / read_restart.cpp,
lmprestart.h), so files move between this package, the reference and LAMMPS:

    magic "LammpS RestartT\\0", int endian (1), int format revision (3)
    header        (int flag, payload) pairs up to flag -1
    groups        int count, then length-prefixed names
    type arrays   (int flag, payload) pairs up to flag -1
    force field   (int flag, payload) pairs up to flag -1
    fix state     counted list of (id, style, blob), then counted list of
                  (id, style, max per-atom size) for fixes with per-atom data
    file layout   (int flag, payload) pairs up to flag -1
    PERPROC chunks: int PERPROC, int ndoubles, ndoubles little-endian doubles
    magic again

Each atom record in a chunk is
    [len, x, y, z, tag, type, mask, image, vx, vy, vz, radius, rmass,
     omegax, omegay, omegaz, <per-atom fix data>]
with tag/type/mask/image int64 bit patterns stored in the doubles. The fix
data of a granular run is FixNeighHistory's record
    [len, npartner, tag_0, shear_0[3], tag_1, shear_1[3], ...]
holding every contact under both atoms, the second copy negated.

Atom records are packed and unpacked with numpy over whole arrays; only
atoms whose records differ in length (contact history) are walked one by one
on reading.
"""

from __future__ import annotations

import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

MAGIC = b"LammpS RestartT\x00"
# read_restart passes VERSION through utils::date2num(): it must be a LAMMPS
# release date. The layout below is that release's.
VERSION_STRING = "2 Sep 2026"
ENDIAN = 1
FORMAT_REVISION = 3

# lmprestart.h header flags, numbered in declaration order.
_FLAG_NAMES = (
    "VERSION SMALLINT TAGINT BIGINT UNITS NTIMESTEP DIMENSION NPROCS PROCGRID "
    "NEWTON_PAIR NEWTON_BOND XPERIODIC YPERIODIC ZPERIODIC BOUNDARY ATOM_STYLE NATOMS "
    "NTYPES NBONDS NBONDTYPES BOND_PER_ATOM NANGLES NANGLETYPES ANGLE_PER_ATOM NDIHEDRALS "
    "NDIHEDRALTYPES DIHEDRAL_PER_ATOM NIMPROPERS NIMPROPERTYPES IMPROPER_PER_ATOM "
    "TRICLINIC BOXLO BOXHI XY XZ YZ SPECIAL_LJ SPECIAL_COUL MASS PAIR BOND ANGLE DIHEDRAL "
    "IMPROPER MULTIPROC MPIIO PROCSPERFILE PERPROC IMAGEINT BOUNDMIN TIMESTEP ATOM_ID "
    "ATOM_MAP_STYLE ATOM_MAP_USER ATOM_SORTFREQ ATOM_SORTBIN COMM_MODE COMM_CUTOFF COMM_VEL "
    "NO_PAIR EXTRA_BOND_PER_ATOM EXTRA_ANGLE_PER_ATOM EXTRA_DIHEDRAL_PER_ATOM "
    "EXTRA_IMPROPER_PER_ATOM EXTRA_SPECIAL_PER_ATOM ATOM_MAXSPECIAL NELLIPSOIDS NLINES "
    "NTRIS NBODIES ATIME ATIMESTEP LABELMAP TRICLINIC_GENERAL ROTATE_G2R ATOM_MAXEXCHANGE"
).split()
FLAG = {name: k for k, name in enumerate(_FLAG_NAMES)}
PERPROC = FLAG["PERPROC"]

# payload kind of each header flag
_KIND = {}
for _names, _kind in (
    ("SMALLINT TAGINT BIGINT IMAGEINT DIMENSION NPROCS NEWTON_PAIR NEWTON_BOND XPERIODIC "
     "YPERIODIC ZPERIODIC NTYPES NBONDTYPES BOND_PER_ATOM NANGLETYPES ANGLE_PER_ATOM "
     "NDIHEDRALTYPES DIHEDRAL_PER_ATOM NIMPROPERTYPES IMPROPER_PER_ATOM TRICLINIC "
     "TRICLINIC_GENERAL MULTIPROC MPIIO PROCSPERFILE ATOM_ID ATOM_MAP_STYLE ATOM_MAP_USER "
     "ATOM_SORTFREQ COMM_MODE COMM_VEL NO_PAIR EXTRA_BOND_PER_ATOM EXTRA_ANGLE_PER_ATOM "
     "EXTRA_DIHEDRAL_PER_ATOM EXTRA_IMPROPER_PER_ATOM EXTRA_SPECIAL_PER_ATOM ATOM_MAXSPECIAL "
     "ATOM_MAXEXCHANGE LABELMAP", "int"),
    ("NTIMESTEP NATOMS NBONDS NANGLES NDIHEDRALS NIMPROPERS NELLIPSOIDS NLINES NTRIS "
     "NBODIES ATIMESTEP", "bigint"),
    ("TIMESTEP XY XZ YZ ATIME ATOM_SORTBIN COMM_CUTOFF", "double"),
    ("VERSION UNITS ATOM_STYLE", "string"),
    ("PROCGRID BOUNDARY", "int_vec"),
    ("BOXLO BOXHI SPECIAL_LJ SPECIAL_COUL BOUNDMIN ROTATE_G2R MASS", "double_vec"),
    ("PAIR NO_PAIR", "pair"),
):
    for _n in _names.split():
        _KIND[FLAG[_n]] = _kind

_BOUNDARY_CHAR = {0: "p", 1: "f", 2: "s", 3: "m"}
# GranularModel::write_restart: one (name, coefficients) entry per slot
_GRAN_SLOTS = ("normal", "damping", "tangential", "rolling", "twisting", "heat")
_LEGACY_STYLES = ("gran/hooke", "gran/hooke/history", "gran/hertz/history")
# FixNeighHistory's id per pair style (PairGranular::init_style,
# PairGranHookeHistory::set_history_id, instance 0): LAMMPS matches the
# per-atom restart data on id and style.
_HISTORY_FIX_ID = {
    "granular": "NEIGH_HISTORY_GRANULAR0",
    "gran/hooke/history": "NEIGH_HISTORY_HH0",
    "gran/hertz/history": "NEIGH_HISTORY_HH0",
}
# image flags (512, 512, 512) in three 10-bit fields: no boundary crossed
_IMAGE_ZERO = (512 << 20) | (512 << 10) | 512
# doubles of an atom_style sphere record before the fix data
_SPHERE_FIELDS = 16


@dataclass
class RestartData:
    """What a restart file holds that the DEM path can use."""

    timestep: int = 0
    dt: float = 0.0
    natoms: int = 0
    ntypes: int = 1
    dimension: int = 3
    units: str = "si"
    atom_style: str = "sphere"
    boxlo: list[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    boxhi: list[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    tilt: list[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    boundary: tuple[str, str, str] = ("p", "p", "p")
    pair_style: str = ""
    pair_settings: dict = field(default_factory=dict)
    fix_global: dict = field(default_factory=dict)
    fix_peratom_styles: list = field(default_factory=list)
    tag: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    atom_type: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    x: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    v: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    omega: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    radius: np.ndarray = field(default_factory=lambda: np.zeros(0))
    rmass: np.ndarray = field(default_factory=lambda: np.zeros(0))
    # per atom, the doubles after the sphere fields (fix data)
    fix_extra: list[np.ndarray] = field(default_factory=list)

    @property
    def density(self) -> np.ndarray:
        """rmass / (4/3 pi r^3), 0 where the volume is 0."""
        vol = (4.0 / 3.0) * np.pi * self.radius**3
        return np.divide(self.rmass, vol, out=np.zeros_like(self.rmass), where=vol > 0)


# ------------------------------------------------------------------ bit casts

def _bits_to_int(d: np.ndarray) -> np.ndarray:
    """int64 values carried in doubles' bit patterns (LAMMPS ubuf)."""
    return np.ascontiguousarray(d, dtype="<f8").view("<i8")


def _int_to_bits(i) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(i, dtype="<i8")).view("<f8")


# ------------------------------------------------------------------- reading

class _Cursor:
    def __init__(self, raw: bytes) -> None:
        self.raw = raw
        self.pos = 0

    def take(self, n: int) -> bytes:
        if n < 0 or self.pos + n > len(self.raw):
            raise ValueError("restart file ended mid-record")
        b = self.raw[self.pos:self.pos + n]
        self.pos += n
        return b

    def i32(self) -> int:
        return struct.unpack("<i", self.take(4))[0]

    def i64(self) -> int:
        return struct.unpack("<q", self.take(8))[0]

    def f64(self) -> float:
        return struct.unpack("<d", self.take(8))[0]

    def string(self) -> str:
        return self.take(self.i32()).decode("utf-8").rstrip("\x00")

    def ivec(self) -> list[int]:
        n = self.i32()
        return list(struct.unpack(f"<{n}i", self.take(4 * n)))

    def dvec(self) -> list[float]:
        n = self.i32()
        return list(struct.unpack(f"<{n}d", self.take(8 * n)))

    def doubles(self, n: int) -> np.ndarray:
        return np.frombuffer(self.take(8 * n), dtype="<f8").astype(np.float64)


def read_restart(path: str | Path) -> RestartData:
    """A single-file LAMMPS restart of an atom_style sphere system.

    ValueError for a file that is not a restart, has the other byte order, a
    newer format revision, is split over processors (MULTIPROC), holds a pair
    style whose payload layout is unknown, or is cut short.
    """
    c = _Cursor(Path(path).read_bytes())
    if c.take(len(MAGIC)) != MAGIC:
        raise ValueError(f"{path} is not a LAMMPS restart file")
    if c.i32() != ENDIAN:
        raise ValueError(f"{path} was written on a machine with the opposite byte order")
    rev = c.i32()
    if rev > FORMAT_REVISION:
        raise ValueError(f"{path} uses restart format revision {rev}; this reader "
                         f"understands up to {FORMAT_REVISION}")
    d = RestartData()
    _read_flags(c, d)  # header
    ngroup = c.i32()   # groups: deleted groups are stored as empty names
    found = 0
    while found < ngroup:
        k = c.i32()
        if k:
            c.take(k)
            found += 1
    _read_flags(c, d)  # type arrays
    _read_flags(c, d)  # force field
    for _ in range(c.i32()):  # fixes with global state
        c.take(c.i32())
        style = c.take(c.i32()).rstrip(b"\x00").decode("utf-8")
        d.fix_global[style] = c.take(c.i32())
    for _ in range(c.i32()):  # fixes with per-atom state
        c.take(c.i32())
        d.fix_peratom_styles.append(c.take(c.i32()).rstrip(b"\x00").decode("utf-8"))
        c.i32()
    _read_flags(c, d)  # file layout
    _read_atoms(c, d)
    return d


def _read_flags(c: _Cursor, d: RestartData) -> None:
    while True:
        flag = c.i32()
        if flag < 0:
            return
        kind = _KIND.get(flag)
        if kind == "int":
            v = c.i32()
            if flag == FLAG["DIMENSION"]:
                d.dimension = v
            elif flag == FLAG["NTYPES"]:
                d.ntypes = v
            elif flag == FLAG["MULTIPROC"] and v:
                raise ValueError("multi-processor restart files are not supported; write a "
                                 "single-file restart instead")
        elif kind == "bigint":
            v = c.i64()
            if flag == FLAG["NTIMESTEP"]:
                d.timestep = v
            elif flag == FLAG["NATOMS"]:
                d.natoms = v
        elif kind == "double":
            v = c.f64()
            if flag == FLAG["TIMESTEP"]:
                d.dt = v
            elif flag in (FLAG["XY"], FLAG["XZ"], FLAG["YZ"]):
                d.tilt[flag - FLAG["XY"]] = v
        elif kind == "string":
            v = c.string()
            if flag == FLAG["UNITS"]:
                d.units = v
            elif flag == FLAG["ATOM_STYLE"]:
                d.atom_style = v.split()[0]
                for _ in range(c.i32()):  # the style's own arguments, raw
                    c.string()
        elif kind == "int_vec":
            v = c.ivec()
            if flag == FLAG["BOUNDARY"]:
                d.boundary = tuple(_BOUNDARY_CHAR.get(v[2 * k], "f") for k in range(3))
        elif kind == "double_vec":
            v = c.dvec()
            if flag == FLAG["BOXLO"]:
                d.boxlo = v
            elif flag == FLAG["BOXHI"]:
                d.boxhi = v
        elif kind == "pair":
            d.pair_style = c.string()
            if flag == FLAG["PAIR"]:
                _read_pair(c, d)
        else:
            raise ValueError(f"unknown restart header flag {flag}; the file may have been "
                             "written by an incompatible LAMMPS version")


def _read_pair(c: _Cursor, d: RestartData) -> None:
    """The pair style's payload (no flags inside: its bytes must be known)."""
    style = d.pair_style
    if style == "granular":
        models = []
        for _ in range(c.i32()):
            m = {}
            for slot in _GRAN_SLOTS:
                name = c.take(c.i32()).decode("utf-8")
                m[slot] = (name, [c.f64() for _ in range(c.i32())])
            m["limit_damping"] = c.i32()
            models.append(m)
        for i in range(d.ntypes):  # setflag, cutoff, model index per type pair
            for _ in range(i, d.ntypes):
                if c.i32():
                    c.f64()
                    c.i32()
        d.pair_settings = {"models": models}
    elif style in _LEGACY_STYLES:
        s = {k: c.f64() for k in ("kn", "kt", "gamman", "gammat", "xmu")}
        s["dampflag"] = c.i32()
        s["limit_damping"] = c.i32()
        for i in range(d.ntypes):
            for _ in range(i, d.ntypes):
                c.i32()
        d.pair_settings = s
    else:
        raise ValueError(f"restart file stores pair_style {style!r}, whose binary layout "
                         "this reader does not know; its bytes cannot be skipped safely")


def _read_atoms(c: _Cursor, d: RestartData) -> None:
    chunks = []
    # the file ends with a second magic string after the last chunk
    while c.pos + len(MAGIC) <= len(c.raw) and c.raw[c.pos:c.pos + len(MAGIC)] != MAGIC:
        flag = c.i32()
        if flag != PERPROC:
            raise ValueError(f"expected a PERPROC chunk, got flag {flag}")
        chunks.append(c.doubles(c.i32()))
    buf = np.concatenate(chunks) if chunks else np.zeros(0)
    # record offsets: fixed-length records are found by a reshape, mixed ones
    # by walking their length fields
    starts = []
    if len(buf):
        first = int(buf[0])
        if first >= 1 and len(buf) % first == 0 and np.all(buf[::first] == first):
            starts = np.arange(0, len(buf), first)
            sizes = np.full(len(starts), first)
        else:
            m = 0
            while m < len(buf):
                starts.append(m)
                m += max(int(buf[m]), 1)
            starts = np.asarray(starts)
            sizes = buf[starts].astype(np.int64)
        if np.any(sizes < _SPHERE_FIELDS) or starts[-1] + sizes[-1] > len(buf):
            raise ValueError("an atom record is too short for atom_style sphere")
    else:
        sizes = np.zeros(0, dtype=np.int64)
    starts = np.asarray(starts, dtype=np.int64)
    col = lambda k: buf[starts + k] if len(starts) else np.zeros(0)  # noqa: E731
    d.x = np.stack([col(1), col(2), col(3)], axis=1) if len(starts) else np.zeros((0, 3))
    d.tag = _bits_to_int(col(4)).astype(np.int32)
    d.atom_type = _bits_to_int(col(5)).astype(np.int32)
    d.v = np.stack([col(8), col(9), col(10)], axis=1) if len(starts) else np.zeros((0, 3))
    d.radius = col(11).copy()
    d.rmass = col(12).copy()
    d.omega = np.stack([col(13), col(14), col(15)], axis=1) if len(starts) else np.zeros((0, 3))
    d.fix_extra = [buf[s + _SPHERE_FIELDS:s + z].copy() for s, z in zip(starts, sizes)]
    if d.natoms and len(d.tag) != d.natoms:
        raise ValueError(f"restart header declares {d.natoms} atoms but {len(d.tag)} "
                         "were found")


def parse_neigh_history(record: np.ndarray, dnum: int = 3) -> tuple[np.ndarray, np.ndarray]:
    """One FixNeighHistory per-atom record [len, npartner, (tag, value[dnum])...]
    as (partner tags int32[npartner], values [npartner, dnum])."""
    record = np.asarray(record, dtype=np.float64)
    if record.size == 0:
        return np.zeros(0, dtype=np.int32), np.zeros((0, dnum))
    npartner = int(record[1])
    body = record[2:2 + npartner * (1 + dnum)].reshape(npartner, 1 + dnum)
    return _bits_to_int(body[:, 0]).astype(np.int32), body[:, 1:].copy()


# ------------------------------------------------------------------- writing

class _Out:
    def __init__(self) -> None:
        self.parts: list[bytes] = []

    def raw(self, b: bytes) -> None:
        self.parts.append(b)

    def i32(self, *v: int) -> None:
        self.parts.append(struct.pack(f"<{len(v)}i", *v))

    def counted(self, text: str, nul: bool = True) -> None:
        b = text.encode("utf-8") + (b"\x00" if nul else b"")
        self.parts.append(struct.pack("<i", len(b)) + b)

    def flag(self, name: str, value) -> None:
        f = FLAG[name]
        kind = _KIND[f]
        if kind == "int":
            self.parts.append(struct.pack("<ii", f, int(value)))
        elif kind == "bigint":
            self.parts.append(struct.pack("<iq", f, int(value)))
        elif kind == "double":
            self.parts.append(struct.pack("<id", f, float(value)))
        elif kind in ("string", "pair"):
            self.i32(f)
            self.counted(value)
        elif kind == "int_vec":
            self.parts.append(struct.pack(f"<ii{len(value)}i", f, len(value), *value))
        else:
            self.parts.append(struct.pack(f"<ii{len(value)}d", f, len(value), *value))

    def end(self) -> None:
        self.i32(-1)


def write_restart(path: str | Path, atom, domain, timestep: int, dt: float, history=None,
                  units: str = "si", pair=None, pair_style: str = "granular",
                  nlist=None) -> None:
    """Write a single-file restart of `atom` in `domain` (write_restart command).

    With `history` and `nlist` the touching contacts' shear history is stored
    in the FixNeighHistory layout, each contact under both atoms (the second
    copy negated), as LAMMPS stores it. With `pair` (a PairGranular) the force
    field section holds its sub-models; other pair styles raise ValueError.
    """
    n = atom.nlocal
    o = _Out()
    o.raw(MAGIC)
    o.i32(ENDIAN, FORMAT_REVISION)
    # header
    o.flag("VERSION", VERSION_STRING)
    o.flag("SMALLINT", 4)
    o.flag("IMAGEINT", 4)
    o.flag("TAGINT", 4)
    o.flag("BIGINT", 8)
    o.flag("UNITS", units)
    o.flag("NTIMESTEP", timestep)
    o.flag("DIMENSION", 3)
    o.flag("NPROCS", 1)
    o.flag("PROCGRID", [1, 1, 1])
    # every pair is evaluated once with both halves applied: newton off
    o.flag("NEWTON_PAIR", 0)
    o.flag("NEWTON_BOND", 0)
    pull = getattr(domain, "pull", None)
    if pull is not None:
        pull()
    per = [int(p) for p in domain.periodicity]
    o.flag("XPERIODIC", per[0])
    o.flag("YPERIODIC", per[1])
    o.flag("ZPERIODIC", per[2])
    o.flag("BOUNDARY", [0 if per[k] else 1 for k in range(3) for _ in range(2)])
    o.flag("ATOM_STYLE", "sphere")
    o.i32(0)  # atom_style sphere has no sub-style arguments
    o.flag("NATOMS", n)
    atype = atom.atom_type.to_numpy()[:n]
    ntypes = int(atype.max()) if n else 1
    o.flag("NTYPES", ntypes)
    o.flag("TRICLINIC", 1 if domain.triclinic else 0)
    o.flag("BOXLO", [float(v) for v in domain.boxlo])
    o.flag("BOXHI", [float(v) for v in domain.boxhi])
    if domain.triclinic:
        for name, val in zip(("XY", "XZ", "YZ"), domain.tilt):
            o.flag(name, val)
    o.flag("TIMESTEP", dt)
    o.end()
    # groups: "all"
    o.i32(1)
    o.counted("all")
    # type arrays: none (per-atom mass)
    o.end()
    # force field
    if pair is not None:
        _write_pair(o, pair, pair_style, ntypes)
    o.end()
    # fix state
    tag = atom.tag.to_numpy()[:n].astype(np.int64)
    hist = (_history_records(history, nlist, tag)
            if history is not None and nlist is not None else None)
    o.i32(0)
    if hist is not None:
        o.i32(1)
        o.counted(_HISTORY_FIX_ID.get(pair_style, _HISTORY_FIX_ID["granular"]))
        o.counted("NEIGH_HISTORY")
        o.i32(int(2 + 4 * hist[0].max()) if n and hist[0].max() > 0 else 2)
    else:
        o.i32(0)
    # file layout
    o.flag("MULTIPROC", 0)
    o.end()
    payload = _pack_atoms(atom, n, tag, atype, hist)
    o.i32(PERPROC, len(payload))
    o.raw(payload.astype("<f8").tobytes())
    o.raw(MAGIC)
    Path(path).write_bytes(b"".join(o.parts))


def _write_pair(o: _Out, pair, style: str, ntypes: int) -> None:
    sub = getattr(pair, "submodels", None)
    if style != "granular" or sub is None:
        raise ValueError(f"storing pair_style {style!r} in a restart file is not implemented; "
                         "omit `pair` to write the file without a force field section")
    o.flag("PAIR", "granular")
    o.i32(1)  # one model
    for slot in _GRAN_SLOTS:
        name, coeffs = sub[slot]
        o.counted(name, nul=False)
        o.i32(len(coeffs))
        if coeffs:
            o.raw(struct.pack(f"<{len(coeffs)}d", *coeffs))
    o.i32(int(getattr(pair, "limit_damping", 0)))
    # setflag 1, cutoff -1 (derived from the radii), model 0 per type pair
    for i in range(ntypes):
        for _ in range(i, ntypes):
            o.raw(struct.pack("<idi", 1, -1.0, 0))


def _history_records(history, nlist, tag: np.ndarray):
    """Every touching contact under both atoms, as LAMMPS stores it: (count[n],
    then per entry its atom, partner tag and shear, in write order). A
    contact (i, partner) of the half list appears under i as stored and under
    the partner negated, the shear being measured from i towards j. Entries
    are in the order of a walk over atoms i and their contacts in pair order,
    each contact written to i and then to its partner."""
    n = len(tag)
    m = int(nlist.npairs[None])
    pi = nlist.pair_i.to_numpy()[:m].astype(np.int64)
    partner = history.partner.to_numpy()[:m].astype(np.int64)
    shear = (np.asarray(history.shear.to_numpy()[:m], dtype=np.float64).reshape(m, -1)[:, :3]
             if m else np.zeros((0, 3)))
    nc = np.arange(m)
    on = (partner >= 0) & (pi < n)
    pi, partner, shear, nc = pi[on], partner[on], shear[on], nc[on]
    # the partner's local index from its tag
    order = np.argsort(tag, kind="stable")
    pos = np.minimum(np.searchsorted(tag[order], partner), max(n - 1, 0))
    found = tag[order][pos] == partner if n else np.zeros(0, dtype=bool)
    pj = order[pos]
    dest = np.concatenate([pi, pj[found]])
    walk = np.concatenate([pi, pi[found]])
    step = np.concatenate([2 * nc, 2 * nc[found] + 1])
    ptag = np.concatenate([partner, tag[pi[found]]])
    val = np.concatenate([shear, -shear[found]])
    w = np.lexsort((step, walk, dest))
    return np.bincount(dest, minlength=n), dest[w], ptag[w], val[w]


def _pack_atoms(atom, n: int, tag: np.ndarray, atype: np.ndarray, hist) -> np.ndarray:
    x = atom.x.to_numpy()[:n]
    v = atom.v.to_numpy()[:n]
    omega = atom.omega.to_numpy()[:n]
    radius = atom.radius.to_numpy()[:n]
    rmass = atom.rmass.to_numpy()[:n]
    base = np.empty((n, _SPHERE_FIELDS))
    base[:, 1:4] = x
    base[:, 4] = _int_to_bits(tag)
    base[:, 5] = _int_to_bits(atype.astype(np.int64))
    base[:, 6] = _int_to_bits(np.ones(n, dtype=np.int64))
    base[:, 7] = _int_to_bits(np.full(n, _IMAGE_ZERO, dtype=np.int64))
    base[:, 8:11] = v
    base[:, 11] = radius
    base[:, 12] = rmass
    base[:, 13:16] = omega
    if hist is None:
        base[:, 0] = _SPHERE_FIELDS
        return base.reshape(-1)
    count, owner, ptag, val = hist
    hlen = 2 + 4 * count
    size = _SPHERE_FIELDS + hlen
    base[:, 0] = size
    start = np.concatenate([[0], np.cumsum(size)[:-1]]).astype(np.int64)
    out = np.empty(int(size.sum()))
    cols = start[:, None] + np.arange(_SPHERE_FIELDS)
    out[cols] = base
    out[start + _SPHERE_FIELDS] = hlen
    out[start + _SPHERE_FIELDS + 1] = count
    # entry k of atom a sits at start[a] + 18 + 4 k
    first = np.concatenate([[0], np.cumsum(count)[:-1]]).astype(np.int64)
    k = np.arange(len(owner)) - first[owner]
    at = start[owner] + _SPHERE_FIELDS + 2 + 4 * k
    out[at] = _int_to_bits(ptag)
    out[at + 1] = val[:, 0]
    out[at + 2] = val[:, 1]
    out[at + 3] = val[:, 2]
    return out
