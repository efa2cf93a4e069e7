"""LAMMPS input scripts on the device DEM path (reference: input.py, main.py).

`LAMMPSInputParser(script).execute()` runs a granular LAMMPS script through
this package's Simulation, so a run is the device-resident tmps_run whenever
its fixes allow. The commands of the DEM path are understood:

    units atom_style dimension boundary newton comm_modify atom_modify
    processors log echo thermo_modify       (accepted; settings of no effect)
    variable <name> equal|string|index ...  print "<text>"
    region <id> block|cylinder ...   create_box <n> <region>   read_data <file>
    read_restart <file>     write_restart <file>     restart <N> <file>
    pair_style gran/hooke/history|gran/hertz/history Kn Kt gn gt xmu damp [limit_damping]
    pair_style granular  +  pair_coeff * * <normal> <coeffs> [tangential ..] [damping ..]
    neighbor <skin> bin     neigh_modify delay|every|check ...
    timestep <dt>           thermo <N>   thermo_style custom <keywords>
    fix <id> <group> nve/sphere | damping/cundall gl ga |
        deform[/pressure] N <dim> <style> <args> ... [couple|remap|max/rate ...] |
        wall/gran hooke[/history] Kn Kt gn gt xmu damp xplane|yplane|zplane lo hi |
        wall/gran[/region] hooke|hertz[/history] ... zcylinder|cylinder ...|region <cyl> |
        print N "<text with ${var}>" [file <path>] [screen yes|no] [title <text>]
    unfix <id>
    compute <id> <group> contact/atom | pressure <temp|NULL> | temp | ke |
        erotate/sphere | stress/atom <temp|NULL> | fabric | reduce <mode> c_<id>[i] ...
    run <N>

Expressions are arithmetic (+ - * / ^, parentheses, sqrt exp ln log sin cos
tan abs floor ceil, PI) over numbers, ${name} / $x substitutions, v_name and
c_id / c_id[i] references. Commands outside the path (dump, fix gravity,
region walls, ...) raise NotImplementedError naming them.
"""

from __future__ import annotations

import ast
import math
import operator
import re
import shlex
from pathlib import Path
from typing import Any

import numpy as np

from . import restart as RS
from .atom import AtomSystem
from .computes import ComputeContactAtom, ComputeFabric, Computes, ComputeStressAtom
from .constants import f64
from .contact_history import ContactHistory
from .data_reader import read_data
from .domain import Domain
from .EXTRA_FIX import (FixDampingCundall, FixDeformPressure, FixNVESphere, FixPrint,
                        FixWallGran, FixWallGranRegion)
from .GRANULAR import GranHertzHistory, GranHookeHistory, PairGranular
from .neighbor import NeighborList
from .simulation import Simulation

_IGNORED = {"units", "dimension", "newton", "comm_modify", "atom_modify", "processors", "log",
            "echo", "thermo_modify", "comm_style", "neigh_modify_off", "clear"}
_OUT_OF_PATH = {"dump", "undump", "dump_modify", "create_atoms", "lattice", "velocity",
                "set", "group", "displace_atoms", "delete_atoms", "minimize", "jump",
                "label", "next", "include", "shell"}

# granular sub-model coefficient counts (PairGranular::coeff)
GRANULAR_NCOEFF = {
    "normal": {"hooke": 2, "hertz": 2, "hertz/material": 3},
    "tangential": {"none": 0, "linear_nohistory": 2, "linear_history": 3, "mindlin": 3},
    "damping": {"none": 0, "velocity": 0, "mass_velocity": 0, "viscoelastic": 0, "tsuji": 0,
                "coeff_restitution": 0},
}
# values each fix deform style takes after its name, for a box dimension
DEFORM_NARG = {"final": 2, "delta": 2, "scale": 1, "vel": 1, "erate": 1, "trate": 1,
               "volume": 0, "wiggle": 2, "variable": 2, "pressure": 2, "pressure/mean": 2}

_BINOPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
           ast.Div: operator.truediv, ast.Pow: operator.pow, ast.Mod: operator.mod}
_FUNCS = {"sqrt": math.sqrt, "exp": math.exp, "ln": math.log, "log": math.log10,
          "sin": math.sin, "cos": math.cos, "tan": math.tan, "abs": abs, "floor": math.floor,
          "ceil": math.ceil, "atan": math.atan, "asin": math.asin, "acos": math.acos}


# ------------------------------------------------------------ pure parsing

def parse_gran_settings(args: list[str], ev) -> dict:
    """`Kn Kt gamma_n gamma_t xmu dampflag [limit_damping]` with the NULL
    shorthands Kt = 2/7 Kn, gamma_t = gamma_n / 2 (input.py:444-472)."""
    if len(args) < 6:
        raise ValueError("pair_style gran/* needs Kn Kt gamma_n gamma_t xmu dampflag")
    kn = ev(args[0])
    kt = kn * 2.0 / 7.0 if args[1] == "NULL" else ev(args[1])
    gamman = ev(args[2])
    gammat = 0.5 * gamman if args[3] == "NULL" else ev(args[3])
    out = dict(kn=kn, kt=kt, gamman=gamman, gammat=gammat, xmu=ev(args[4]),
               dampflag=int(ev(args[5])), limit_damping=False)
    if len(args) > 6:
        if args[6] != "limit_damping":
            raise ValueError(f"Unknown pair_style gran/* keyword: {args[6]}")
        out["limit_damping"] = True
    return out


def parse_pair_coeff_granular(args: list[str], ev) -> dict:
    """`pair_coeff i j <normal> <coeffs> [tangential <name> <coeffs>]
    [damping <name>] [limit_damping]` as PairGranular keyword arguments;
    damping defaults to viscoelastic (input.py:380-442)."""
    if len(args) < 3:
        raise ValueError("pair_coeff granular needs at least a normal model")

    def take(name, category, i):
        table = GRANULAR_NCOEFF[category]
        if name not in table:
            raise ValueError(f"Unsupported pair granular {category} model {name!r}; "
                             f"implemented: {sorted(table)}")
        n = table[name]
        if i + n > len(args):
            raise ValueError(f"pair_coeff granular: {name} needs {n} coefficients")
        return [None if t == "NULL" else ev(t) for t in args[i:i + n]], i + n

    normal = args[2]
    ncoef, i = take(normal, "normal", 3)
    out = dict(normal=normal, normal_coeffs=[c for c in ncoef if c is not None],
               tangential="none", tangential_coeffs=[], damping="viscoelastic",
               limit_damping=False)
    while i < len(args):
        key = args[i]
        if key == "tangential":
            out["tangential"] = args[i + 1]
            out["tangential_coeffs"], i = take(args[i + 1], "tangential", i + 2)
        elif key == "damping":
            out["damping"] = args[i + 1]
            _, i = take(args[i + 1], "damping", i + 2)
        elif key == "limit_damping":
            out["limit_damping"] = True
            i += 1
        elif key in ("rolling", "twisting"):
            raise NotImplementedError(f"pair granular {key} resistance is not implemented "
                                      "(the reference raises on it too)")
        else:
            raise ValueError(f"Unsupported pair_coeff granular keyword: {key!r}")
    return out


def parse_deform(args: list[str], ev) -> dict:
    """`N <dim> <style> <values> ... [keywords]` as FixDeformPressure keyword
    arguments (input.py:287-378); box dimensions x, y, z."""
    if not args:
        raise ValueError("fix deform needs an N argument")
    out: dict[str, Any] = dict(nevery=int(ev(args[0])), axes={}, box=None, couple="none",
                               max_rate=0.0, normalize_pressure=False, vol_balance_p=False,
                               remap="x", flip=True)
    i = 1
    while i < len(args):
        tok = args[i]
        if tok in ("x", "y", "z"):
            style = args[i + 1]
            if style not in DEFORM_NARG:
                raise ValueError(f"Unsupported fix deform style {style!r} for {tok}")
            n = DEFORM_NARG[style]
            vals = args[i + 2:i + 2 + n]
            spec: dict[str, Any] = {"style": style}
            if style == "variable":
                spec["hstr"], spec["hratestr"] = (v.removeprefix("v_") for v in vals)
            else:
                nums = [ev(v) for v in vals]
                if style == "final":
                    spec["flo"], spec["fhi"] = nums
                elif style == "delta":
                    spec["dlo"], spec["dhi"] = nums
                elif style == "scale":
                    spec["scale"] = nums[0]
                elif style == "vel":
                    spec["vel"] = nums[0]
                elif style in ("erate", "trate"):
                    spec["rate"] = nums[0]
                elif style == "wiggle":
                    spec["amplitude"], spec["tperiod"] = nums
                elif style in ("pressure", "pressure/mean"):
                    spec["ptarget"], spec["pgain"] = nums
            out["axes"][tok] = spec
            i += 2 + n
        elif tok in ("xy", "xz", "yz"):
            raise NotImplementedError("tilt factors are not on the device DEM path")
        elif tok == "box":
            style = args[i + 1]
            if style == "volume":
                out["box"] = {"style": "volume"}
                i += 2
            elif style == "pressure":
                out["box"] = {"style": "pressure", "ptarget": ev(args[i + 2]),
                              "pgain": ev(args[i + 3])}
                i += 4
            else:
                raise ValueError(f"Unsupported fix deform box style {style!r}")
        elif tok == "couple":
            out["couple"] = args[i + 1]
            i += 2
        elif tok == "max/rate":
            out["max_rate"] = ev(args[i + 1])
            i += 2
        elif tok in ("normalize/pressure", "vol/balance/p", "flip"):
            key = {"normalize/pressure": "normalize_pressure", "vol/balance/p": "vol_balance_p",
                   "flip": "flip"}[tok]
            out[key] = args[i + 1].lower() == "yes"
            i += 2
        elif tok == "remap":
            out["remap"] = args[i + 1]
            i += 2
        elif tok == "units":
            if args[i + 1].lower() != "box":
                raise ValueError("fix deform units lattice needs a lattice command")
            i += 2
        else:
            raise ValueError(f"Unsupported fix deform keyword: {tok!r}")
    return out


def parse_wall_gran(args: list[str], ev, boxlo, regions=None) -> list[dict]:
    """`<fstyle> Kn Kt gn gt xmu dampflag` then `xplane|yplane|zplane lo hi`
    (FixWallGran, one wall per non-NULL bound: lo side -1, hi side +1; no bound
    means the box's lower face), `zcylinder R [c1 c2]` / `cylinder dim c1 c2 R`
    or `region <id>` of a cylinder region (FixWallGranRegion, marked with
    "cylinder": True); Kt NULL = 2/7 Kn, gamma_t NULL = gamma_n / 2
    (input.py:1454-1640)."""
    i = 0
    fstyle = "hooke"
    params: list[float | None] = []
    if i < len(args) and args[i].lower() in ("hooke", "hooke/history", "hertz",
                                             "hertz/history", "granular"):
        fstyle = args[i].lower()
        i += 1
        while i < len(args) and args[i].lower() not in ("xplane", "yplane", "zplane",
                                                        "cylinder", "zcylinder", "region"):
            if args[i].upper() == "NULL":
                params.append(None)
            else:
                try:
                    params.append(ev(args[i]))
                except ValueError:
                    break  # a sub-model name of the granular style
            i += 1
    if fstyle == "granular":
        raise NotImplementedError("fix wall/gran granular is not on the device DEM path")
    get = (lambda k, d: params[k] if len(params) > k and params[k] is not None else d)
    kn = get(0, 1e5)
    kt, gamman, gammat = get(1, None), get(2, 0.0), get(3, None)
    xmu, dampflag = get(4, 0.0), int(get(5, 1))
    axis, bounds, cyl = None, [None, None], None
    while i < len(args):
        tok = args[i].lower()
        if tok in ("xplane", "yplane", "zplane"):
            axis = "xyz".index(tok[0])
            for slot in range(2):
                if i + 1 + slot < len(args) and args[i + 1 + slot].upper() != "NULL":
                    bounds[slot] = ev(args[i + 1 + slot])
            i += 3
        elif tok == "zcylinder":
            cyl = dict(axis=2, radius=ev(args[i + 1]), c1=0.0, c2=0.0, side="in")
            i += 2
            if i + 1 < len(args) and not args[i].isalpha():
                cyl["c1"], cyl["c2"] = ev(args[i]), ev(args[i + 1])
                i += 2
        elif tok == "cylinder":
            cyl = dict(axis="xyz".index(args[i + 1].lower()), c1=ev(args[i + 2]),
                       c2=ev(args[i + 3]), radius=ev(args[i + 4]), side="in")
            i += 5
        elif tok == "region":
            reg = (regions or {}).get(args[i + 1])
            if not isinstance(reg, dict):
                raise NotImplementedError("fix wall/gran/region needs a cylinder region")
            cyl = dict(reg)
            i += 2
        else:
            i += 1
    if cyl is not None:
        return [dict(cylinder=True, fstyle=fstyle, kn=kn, gamman=gamman, kt=kt, gammat=gammat,
                     xmu=xmu, dampflag=dampflag, **cyl)]
    if fstyle not in ("hooke", "hooke/history"):
        raise NotImplementedError(f"fix wall/gran {fstyle} planes are not on the device DEM path")
    if axis is None:
        axis = 2
    if bounds == [None, None]:
        bounds = [float(boxlo[axis]), None]
    common = dict(wall_axis=axis, kn=kn, gamman=gamman,
                  kt=kt if kt is not None else kn * 2.0 / 7.0,
                  gammat=gammat if gammat is not None else 0.5 * gamman, xmu=xmu,
                  dampflag=dampflag, history=fstyle.endswith("history"))
    return [dict(common, wall_side=-1 if slot == 0 else 1, wall_coord=c)
            for slot, c in enumerate(bounds) if c is not None]


# ------------------------------------------------------------------ parser

class LAMMPSInputParser:
    """Runs a LAMMPS script (reference: input.py LAMMPSInputParser)."""

    def __init__(self, script_path, default_fp: Any = f64, arch: Any = None) -> None:
        self.script_path = Path(script_path)
        self.workdir = self.script_path.parent
        self.default_fp = default_fp
        self.arch = arch
        self.commands: list[str] = []
        self.variables: dict[str, Any] = {}
        self.variable_exprs: dict[str, str] = {}
        self.domain: Domain | None = None
        self.atom: AtomSystem | None = None
        self.neighbor: NeighborList | None = None
        self.history: ContactHistory | None = None
        self.simulation: Simulation | None = None
        self.pair_style = None
        self.pair_style_name: str | None = None
        self.fixes: dict[str, Any] = {}
        self.computes: dict[str, tuple[str, list[str]]] = {}
        self.regions: dict[str, tuple[list[float], list[float]]] = {}
        self.boundary = ("p", "p", "p")
        self.dt = 0.001
        self.neighbor_skin = 0.001
        self.neigh = dict(delay=0, every=1, check=True)
        self.thermo_freq = 0
        self.thermo_keys = ["step", "atoms", "ke", "press"]
        self.thermo_lines: list[str] = []
        self.start_timestep = 0
        self.restart_every = 0
        self.restart_file: str | None = None
        self._seed = None
        self._resolving: set[str] = set()

    # ------------------------------------------------------------- script
    def load_script(self) -> None:
        """Lines with comments removed and `&` continuations joined."""
        text = self.script_path.read_text()
        cmds, buf = [], ""
        for raw in text.splitlines():
            line = _strip_comment(raw).rstrip()
            if line.endswith("&"):
                buf += line[:-1] + " "
                continue
            line = (buf + line).strip()
            buf = ""
            if line:
                cmds.append(line)
        if buf.strip():
            cmds.append(buf.strip())
        self.commands = cmds

    def _subst(self, s: str) -> str:
        def one(m):
            name = m.group(1) or m.group(2)
            if name not in self.variables and name not in self.variable_exprs:
                kw = self._keyword(name)
                if kw is None:
                    raise ValueError(f"substitution of an undefined variable {name!r}")
                return _fmt(kw)
            v = self.evaluate_variable(name)
            return _fmt(v)
        return re.sub(r"\$\{([A-Za-z0-9_]+)\}|\$([A-Za-z0-9_])", one, s)

    def parse_line(self, line: str) -> list[str]:
        line = _strip_comment(line).strip()
        if not line:
            return []
        toks = line.split(None, 1)
        if toks[0] == "print":  # printed text keeps its quotes' content whole
            return ["print", self._subst(toks[1].strip().strip("\"'")) if len(toks) > 1 else ""]
        head = line.split()
        if len(head) >= 4 and head[0] == "fix" and head[3] == "print":
            # fix print's text is substituted when it prints, its other words
            # now (input.py:788-794)
            try:
                words = shlex.split(line, posix=True)
            except ValueError:
                words = head
            return [w if k == 5 else self._subst(w) for k, w in enumerate(words)]
        line = self._subst(line)
        try:
            return shlex.split(line, posix=True)
        except ValueError:
            return line.split()

    # -------------------------------------------------------- expressions
    def evaluate_expression(self, expr: str) -> float:
        expr = self._subst(str(expr).strip().strip("\"'"))
        expr = re.sub(r"c_([A-Za-z0-9_]+)\[(\d+)\]", r"__c('\1',\2)", expr)
        expr = re.sub(r"c_([A-Za-z0-9_]+)", r"__c('\1',0)", expr)
        expr = re.sub(r"v_([A-Za-z0-9_]+)", r"__v('\1')", expr)
        expr = expr.replace("^", "**")
        try:
            tree = ast.parse(expr, mode="eval")
        except SyntaxError as e:
            raise ValueError(f"cannot evaluate expression {expr!r}") from e
        return float(self._eval(tree.body))

    def _eval(self, node):
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return node.value
        if isinstance(node, ast.BinOp) and type(node.op) in _BINOPS:
            return _BINOPS[type(node.op)](self._eval(node.left), self._eval(node.right))
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = self._eval(node.operand)
            return -v if isinstance(node.op, ast.USub) else v
        if isinstance(node, ast.Name):
            if node.id == "PI":
                return math.pi
            # a state keyword first, so `variable lx equal lx` reads the box
            # (the reference refreshes these names before each evaluation)
            kw = self._keyword(node.id)
            if kw is not None:
                return kw
            if node.id in self.variables or node.id in self.variable_exprs:
                return self.evaluate_variable(node.id)
            raise ValueError(f"unknown name {node.id!r} in an expression")
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name):
            name = node.func.id
            if name == "__v":
                return self.evaluate_variable(node.args[0].value)
            if name == "__c":
                return self.compute_value(node.args[0].value, node.args[1].value)
            if name in _FUNCS:
                return _FUNCS[name](*[self._eval(a) for a in node.args])
        raise ValueError(f"unsupported expression element {ast.dump(node)}")

    def evaluate_variable(self, name: str) -> float | str:
        """An equal-style variable is evaluated now (its value is dynamic)."""
        name = name.removeprefix("v_")
        if name in self.variable_exprs:
            if name in self._resolving:
                raise ValueError(f"variable {name!r} refers to itself")
            self._resolving.add(name)
            try:
                return self.evaluate_expression(self.variable_exprs[name])
            finally:
                self._resolving.discard(name)
        if name in self.variables:
            return self.variables[name]
        kw = self._keyword(name)
        if kw is not None:
            return kw
        raise ValueError(f"undefined variable {name!r}")

    def _keyword(self, name: str) -> float | None:
        """The state values an expression may name (input.py:158-198): step,
        lstep, dt, time, lx ly lz (xlength ...), vol, voidratio, ke (translational
        plus rotational), ke_trans, ke_rot, atoms, and e once Vs is stored.
        None for other names."""
        sim, dom, atom = self.simulation, self.domain, self.atom
        step = sim.timestep if sim is not None else self.start_timestep
        if name in ("step", "lstep"):
            return float(step)
        if name == "dt":
            return float(sim.dt if sim is not None else self.dt)
        if name == "time":
            return float(step) * float(sim.dt if sim is not None else self.dt)
        if dom is None:
            return None
        axis = {"lx": 0, "ly": 1, "lz": 2, "xlength": 0, "ylength": 1, "zlength": 2}
        if name in axis:
            dom.pull()
            return float(dom.prd[axis[name]])
        if name == "vol":
            dom.pull()
            return float(dom.volume)
        if name == "e":
            # with Vs (the solid volume) stored, e is the current box's void
            # ratio (input.py:216-221)
            vs = self.variables.get("Vs")
            if isinstance(vs, (int, float)) and not isinstance(vs, bool) and vs > 0.0:
                dom.pull()
                return (float(dom.volume) - float(vs)) / float(vs)
            return None
        if atom is None:
            return None
        if name == "atoms":
            return float(atom.nlocal)
        if name == "voidratio":
            r = atom.radius.to_numpy()[: atom.nlocal]
            solid = float(np.sum(4.0 / 3.0 * np.pi * r ** 3))
            dom.pull()
            return (float(dom.volume) - solid) / solid if solid > 0.0 else 0.0
        if name in ("ke", "ke_trans", "ke_rot"):
            if atom.nlocal == 0:
                return 0.0
            kt, kr = Computes()._ke(atom)
            return float({"ke": kt + kr, "ke_trans": kt, "ke_rot": kr}[name])
        return None

    def evaluate_string_template(self, template_str: str) -> str:
        """fix print's text with ${name} and $name replaced (input.py:115-155):
        an equal-style variable is evaluated now and written with %.15g, and
        one that will not evaluate is an error naming it; a stored number is
        written with %.15g, a string as it is, an unknown name as nothing."""
        def one(m: re.Match) -> str:
            name = m.group(1)
            if name in self.variable_exprs:
                try:
                    v = self.evaluate_variable(name)
                except (ValueError, KeyError, ZeroDivisionError, TypeError) as exc:
                    raise ValueError(f"variable {name!r} could not be evaluated: "
                                     f"{self.variable_exprs[name]!r} ({exc})") from exc
                return f"{float(v):.15g}" if not isinstance(v, str) else v
            kw = self._keyword(name)
            if kw is not None:
                return f"{kw:.15g}"
            v = self.variables.get(name, "")
            if isinstance(v, (int, float)) and not isinstance(v, bool):
                return f"{float(v):.15g}"
            return str(v)

        res = re.sub(r"\$\{([a-zA-Z0-9_]+)\}", one, template_str)
        return re.sub(r"\$([a-zA-Z0-9_]+)", one, res)

    # ------------------------------------------------------------ computes
    def compute_value(self, cid: str, index: int = 0) -> float:
        """c_id (index 0) or c_id[index] (1-based), as a script reads it."""
        if cid not in self.computes:
            raise ValueError(f"unknown compute {cid!r}")
        style, args = self.computes[cid]
        atom, dom = self.atom, self.domain
        if style == "pressure":
            # a NULL temperature compute drops the kinetic term (registry.py:95-115)
            kinetic = args[0].upper() != "NULL"
            dom.pull()
            p = Computes().compute_pressure_tensor(atom, dom, kinetic=kinetic)
            return float(p[:3].mean()) if index == 0 else float(p[index - 1])
        if style == "temp":
            c = Computes()
            return c.temperature(atom) if index == 0 else float(c.kinetic_tensor(atom)[index - 1])
        if style == "ke":
            return Computes().ke_trans(atom)
        if style == "erotate/sphere":
            return Computes().ke_rot(atom)
        if style == "fabric":
            if index == 0:
                return float("nan")  # a vector compute (registry.py:194-203)
            phi, _ = ComputeFabric().compute(atom, dom, self.simulation.neighbor)
            return float(phi[index - 1])
        if style == "reduce":
            mode, inputs = args[0], args[1:]
            if mode not in REDUCE_MODES:
                raise ValueError(f"compute reduce mode {mode!r} is not supported; "
                                 f"implemented: {REDUCE_MODES}")
            src = inputs[max(index, 1) - 1]
            m = re.fullmatch(r"c_([A-Za-z0-9_]+)(?:\[(\d+)\])?", src)
            if m is None:
                raise NotImplementedError("compute reduce over atom variables is not supported")
            vals = np.asarray(self.compute_peratom(m.group(1)), dtype=np.float64)
            if m.group(2) is not None:
                vals = vals[:, int(m.group(2)) - 1]
            return _reduce(mode, vals)
        raise ValueError(f"compute {cid!r} ({style}) has no global value")

    def compute_peratom(self, cid: str) -> np.ndarray:
        style, args = self.computes[cid]
        sim = self.simulation
        if style == "contact/atom":
            return ComputeContactAtom(self.atom.max_atoms).compute(self.atom, self.domain,
                                                                   sim.neighbor)
        if style == "stress/atom":
            kinetic = args[0].upper() != "NULL"
            return ComputeStressAtom(self.atom.max_atoms).compute(self.atom, self.domain,
                                                                  kinetic=kinetic)
        raise ValueError(f"compute {cid!r} ({style}) has no per-atom values")

    # ----------------------------------------------------------- builders
    def _ev(self, tok: str) -> float:
        return self.evaluate_expression(tok)

    def _new_system(self, boxlo, boxhi, natoms: int):
        self.domain = Domain(boxlo, boxhi, self.boundary)
        self.atom = AtomSystem(max_atoms=max(natoms, 1), float_type=self.default_fp)

    def _ensure_simulation(self) -> Simulation:
        if self.simulation is None:
            if self.atom is None or self.domain is None:
                raise ValueError("a run needs atoms: read_data or read_restart first")
            n = self.atom.max_atoms
            self.neighbor = NeighborList(self.domain, skin=self.neighbor_skin, max_atoms=n,
                                         max_neighbors_per_atom=64)
            self.neighbor.delay = self.neigh["delay"]
            self.neighbor.every = self.neigh["every"]
            self.neighbor.check = self.neigh["check"]
            if self.history is None:
                self.history = ContactHistory(n, 64)
            self.simulation = Simulation(self.domain, self.atom, self.neighbor, self.history,
                                         self.pair_style, dt=self.dt)
            self.simulation.timestep = self.start_timestep
            if self.restart_every and self.restart_file:
                self.simulation.set_restart(self.restart_every, self.restart_file,
                                            self._restart_writer)
            for fix in self.fixes.values():
                for one in fix if isinstance(fix, list) else [fix]:
                    self.simulation.add_fix(one)
        return self.simulation

    def _add_fix(self, fid: str, fix) -> None:
        self.fixes[fid] = fix
        if self.simulation is not None:
            for one in fix if isinstance(fix, list) else [fix]:
                self.simulation.add_fix(one)

    def _pair_from(self, name: str, spec: dict):
        if name == "granular":
            return PairGranular(self.domain, **spec)
        cls = {"gran/hooke/history": GranHookeHistory,
               "gran/hertz/history": GranHertzHistory}[name]
        return cls(self.domain, spec["kn"], spec["gamman"], spec["kt"], spec["gammat"],
                   spec["xmu"], dampflag=spec["dampflag"], limit_damping=spec["limit_damping"])

    # -------------------------------------------------------------- thermo
    def thermo_row(self) -> str:
        sim = self.simulation
        self.domain.pull()
        vals = []
        for key in self.thermo_keys:
            if key == "step":
                vals.append(str(sim.timestep if sim else self.start_timestep))
            elif key == "atoms":
                vals.append(str(self.atom.nlocal))
            elif key in ("lx", "ly", "lz"):
                vals.append(_fmt(float(self.domain.prd["xyz".index(key[1])])))
            elif key == "vol":
                vals.append(_fmt(float(self.domain.volume)))
            elif key == "ke":
                vals.append(_fmt(Computes().ke_trans(self.atom)))
            elif key in ("press", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz"):
                p = Computes().compute_pressure_tensor(self.atom, self.domain, kinetic=True)
                idx = {"pxx": 0, "pyy": 1, "pzz": 2, "pxy": 3, "pxz": 4, "pyz": 5}
                vals.append(_fmt(float(p[:3].mean()) if key == "press" else float(p[idx[key]])))
            else:
                vals.append(_fmt(self.evaluate_expression(key)))
        return " ".join(vals)

    # ------------------------------------------------------------- execute
    def execute(self) -> None:
        if not self.commands:
            self.load_script()
        for raw in self.commands:
            tokens = self.parse_line(raw)
            if tokens:
                self.run_command(tokens[0], tokens[1:])

    def run_command(self, cmd: str, args: list[str]) -> None:  # noqa: C901
        ev = self._ev
        if cmd in _IGNORED:
            return
        if cmd in _OUT_OF_PATH:
            raise NotImplementedError(f"{cmd} is not on the device DEM path")
        if cmd == "atom_style":
            if args[0] != "sphere":
                raise ValueError(f"atom_style {args[0]} is not supported; use sphere")
        elif cmd == "variable":
            name, style = args[0], args[1]
            if style == "equal":
                self.variable_exprs[name] = " ".join(args[2:])
                try:
                    self.variables[name] = self.evaluate_expression(self.variable_exprs[name])
                except (ValueError, AttributeError, TypeError, ZeroDivisionError):
                    # refers to something that exists only later (a compute
                    # before the first run): evaluated when it is read
                    self.variables.pop(name, None)
            elif style in ("string", "index"):
                self.variables[name] = args[2]
            else:
                raise NotImplementedError(f"variable style {style} is not supported")
        elif cmd == "print":
            print(args[0], flush=True)
        elif cmd == "boundary":
            self.boundary = tuple(args[:3])
            if self.domain is not None:
                self.domain.set_boundary(self.boundary)
        elif cmd == "region":
            if args[1] == "block":
                b = [ev(t) for t in args[2:8]]
                self.regions[args[0]] = ([b[0], b[2], b[4]], [b[1], b[3], b[5]])
            elif args[1] == "cylinder":
                # region <id> cylinder <dim> <c1> <c2> <radius> <lo> <hi> [side in|out]
                lo_t, hi_t = args[6], args[7]
                bound = (lambda t: None if t.upper() in ("INF", "-INF", "EDGE") else ev(t))
                reg = dict(axis="xyz".index(args[2]), c1=ev(args[3]), c2=ev(args[4]),
                           radius=ev(args[5]), axis_lo=bound(lo_t), axis_hi=bound(hi_t),
                           side="in")
                if "side" in args[8:]:
                    reg["side"] = args[8:][args[8:].index("side") + 1]
                self.regions[args[0]] = reg
            else:
                raise NotImplementedError(f"region style {args[1]} is not supported")
        elif cmd == "create_box":
            reg = self.regions[args[1]]
            if isinstance(reg, dict):
                raise NotImplementedError("create_box needs a block region")
            self._new_system(reg[0], reg[1], 1)
        elif cmd == "read_data":
            d = read_data(self.workdir / args[0])
            if any(d.tilt):
                raise NotImplementedError("triclinic boxes are not on the device DEM path")
            self._new_system(d.boxlo, d.boxhi, d.natoms)
            self.atom.add_particles(x=d.x, radius=d.radius, density=d.density, v=d.v,
                                    omega=d.omega, tag=d.tag, atom_type=d.atom_type)
        elif cmd == "read_restart":
            self._read_restart(self.workdir / args[0])
        elif cmd == "write_restart":
            sim = self.simulation
            step = sim.timestep if sim else self.start_timestep
            self._restart_writer(args[0].replace("*", str(step)), step)
        elif cmd == "restart":
            # restart <N> <file>: N = 0 turns periodic writing off
            self.restart_every = int(ev(args[0]))
            self.restart_file = args[1] if len(args) > 1 else None
            if self.simulation is not None:
                self.simulation.set_restart(self.restart_every, self.restart_file,
                                            self._restart_writer)
        elif cmd == "pair_style":
            self.pair_style_name = args[0]
            if args[0] in ("gran/hooke/history", "gran/hertz/history"):
                self._set_pair(self._pair_from(args[0], parse_gran_settings(args[1:], ev)))
            elif args[0] == "granular":
                self.pair_style = None
            else:
                raise NotImplementedError(f"pair_style {args[0]} is not on the device DEM path")
        elif cmd == "pair_coeff":
            if self.pair_style_name != "granular":
                if len(args) > 2:
                    raise ValueError(f"pair_coeff for pair_style {self.pair_style_name} takes "
                                     "no coefficients; write `pair_coeff * *`")
                return
            self._set_pair(self._pair_from("granular", parse_pair_coeff_granular(args, ev)))
        elif cmd == "neighbor":
            self.neighbor_skin = ev(args[0])
            if self.neighbor is not None:
                self.neighbor.skin = self.neighbor_skin
        elif cmd == "neigh_modify":
            i = 0
            while i < len(args):
                key = args[i]
                if key in ("delay", "every"):
                    self.neigh[key] = int(ev(args[i + 1]))
                    i += 2
                elif key == "check":
                    self.neigh["check"] = args[i + 1].lower() in ("yes", "1", "true")
                    i += 2
                else:
                    i += 1
            if self.neighbor is not None:
                for k, v in self.neigh.items():
                    setattr(self.neighbor, k, v)
        elif cmd == "timestep":
            self.dt = ev(args[0])
            if self.simulation is not None:
                self.simulation.dt = self.dt
        elif cmd == "thermo":
            self.thermo_freq = int(ev(args[0]))
        elif cmd == "thermo_style":
            if args[0] != "custom":
                raise NotImplementedError("only thermo_style custom is supported")
            self.thermo_keys = args[1:]
        elif cmd == "fix":
            self._fix(args[0], args[2], args[3:])
        elif cmd == "unfix":
            fix = self.fixes.pop(args[0], None)
            if fix is not None and self.simulation is not None:
                for one in fix if isinstance(fix, list) else [fix]:
                    self.simulation.remove_fix(one)
        elif cmd == "compute":
            style = args[2]
            if style not in ("contact/atom", "pressure", "temp", "ke", "erotate/sphere", "reduce",
                             "stress/atom", "fabric"):
                raise ValueError(f"Unsupported compute style {style!r}")
            if style in ("pressure", "stress/atom") and len(args) < 4:
                raise ValueError(f"compute {style} needs a temperature compute ID or NULL")
            self.computes[args[0]] = (style, args[3:])
        elif cmd == "run":
            sim = self._ensure_simulation()
            if self.pair_style is not None and sim.pair_style is not self.pair_style:
                sim.set_pair_style(self.pair_style)
            sim.run(int(ev(args[0])))
            if self.thermo_freq:
                row = self.thermo_row()
                self.thermo_lines.append(row)
                print(row, flush=True)
        else:
            raise NotImplementedError(f"command {cmd!r} is not supported")

    def _restart_writer(self, name: str, step: int) -> None:
        """One restart file (write_restart and `restart N`)."""
        sim = self.simulation
        pair = sim.pair_style if sim else self.pair_style
        RS.write_restart(self.workdir / name, self.atom, self.domain, timestep=step, dt=self.dt,
                         history=sim.history if sim else self.history,
                         pair=pair if isinstance(pair, PairGranular) else None,
                         pair_style=self.pair_style_name or "granular",
                         nlist=sim.neighbor if sim else self.neighbor)

    def _set_pair(self, pair) -> None:
        self.pair_style = pair
        if self.simulation is not None:
            self.simulation.set_pair_style(pair)

    def _fix(self, fid: str, style: str, args: list[str]) -> None:
        ev = self._ev
        if self.domain is None:
            raise ValueError(f"fix {style} before the simulation box was defined")
        if style == "nve/sphere":
            self._add_fix(fid, FixNVESphere(self.domain))
        elif style == "damping/cundall":
            self._add_fix(fid, FixDampingCundall(self.domain, ev(args[0]), ev(args[1])))
        elif style in ("deform/pressure", "deform"):
            kw = parse_deform(args, ev)
            self._add_fix(fid, FixDeformPressure(self.domain, var_eval=self.evaluate_variable,
                                                 **kw))
        elif style in ("wall/gran", "wall/gran/region"):
            walls = []
            for kw in parse_wall_gran(args, ev, self.domain.boxlo, self.regions):
                if kw.pop("cylinder", False):
                    walls.append(FixWallGranRegion(self.domain, max_atoms=self.atom.max_atoms,
                                                   **kw))
                else:
                    walls.append(FixWallGran(self.domain, max_atoms=self.atom.max_atoms, **kw))
            self._add_fix(fid, walls[0] if len(walls) == 1 else walls)
        elif style == "print":
            # fix <id> <group> print <N> <text> [file <path>] [screen yes|no]
            # [title <text>] (input.py:1170-1204)
            nevery, text = int(ev(args[0])), args[1]
            path = title = None
            screen = False
            k = 2
            while k < len(args):
                if args[k] in ("file", "title", "screen") and k + 1 < len(args):
                    if args[k] == "file":
                        path = self.workdir / args[k + 1]
                    elif args[k] == "title":
                        title = args[k + 1]
                    else:
                        screen = args[k + 1].lower() == "yes"
                    k += 2
                else:
                    k += 1
            self._add_fix(fid, FixPrint(self.domain, nevery, text, self.evaluate_string_template,
                                        filepath=path, title=title, screen=screen, fix_id=fid))
        else:
            raise NotImplementedError(f"fix {style} is not on the device DEM path")

    def _read_restart(self, path: Path) -> None:
        d = RS.read_restart(path)
        if any(d.tilt):
            raise NotImplementedError("triclinic boxes are not on the device DEM path")
        self.boundary = d.boundary
        self._new_system(d.boxlo, d.boxhi, d.natoms)
        self.atom.add_particles(x=d.x, radius=d.radius, density=d.density, v=d.v,
                                omega=d.omega, tag=d.tag, atom_type=d.atom_type)
        self.atom.rmass.from_numpy(d.rmass)
        self.start_timestep = d.timestep
        if d.dt:
            self.dt = d.dt
        self.history = ContactHistory(self.atom.max_atoms, 64)
        if any(len(e) for e in d.fix_extra):
            dec = [RS.parse_neigh_history(e) for e in d.fix_extra]
            self.history.load_restart([t for t, _ in dec], [v for _, v in dec])
        if d.pair_style == "granular" and d.pair_settings.get("models"):
            m = d.pair_settings["models"][0]
            for slot in ("rolling", "twisting"):
                if m[slot][0] != "none":
                    raise ValueError(f"restart file uses pair granular {slot} model "
                                     f"{m[slot][0]!r}, which is not implemented")
            self.pair_style_name = "granular"
            self.pair_style = PairGranular(
                self.domain, normal=m["normal"][0], normal_coeffs=list(m["normal"][1]),
                tangential=m["tangential"][0],
                tangential_coeffs=[None if abs(c + 1.0) < 1e-15 else c
                                   for c in m["tangential"][1]],
                damping=m["damping"][0], limit_damping=bool(m.get("limit_damping", 0)))
        elif d.pair_style in ("gran/hooke/history", "gran/hertz/history"):
            self.pair_style_name = d.pair_style
            self.pair_style = self._pair_from(d.pair_style, dict(
                d.pair_settings, limit_damping=bool(d.pair_settings["limit_damping"])))


REDUCE_MODES = ("sum", "ave", "min", "max", "sumsq", "avesq", "sumabs", "aveabs", "maxabs")


def _reduce(mode: str, v: np.ndarray) -> float:
    """One compute reduce over per-atom values (registry.py:240-260)."""
    if v.size == 0:
        return 0.0
    op = {"sum": np.sum, "ave": np.mean, "min": np.min, "max": np.max}
    if mode in op:
        return float(op[mode](v))
    if mode in ("sumsq", "avesq"):
        return float((np.sum if mode == "sumsq" else np.mean)(v ** 2))
    if mode in ("sumabs", "aveabs"):
        return float((np.sum if mode == "sumabs" else np.mean)(np.abs(v)))
    return float(np.max(np.abs(v)))


def _strip_comment(line: str) -> str:
    """Text before a # that is not inside quotes."""
    q = None
    for k, ch in enumerate(line):
        if ch in "\"'":
            q = None if q == ch else (ch if q is None else q)
        elif ch == "#" and q is None:
            return line[:k]
    return line


def _fmt(v) -> str:
    if isinstance(v, str):
        return v
    return repr(float(v)) if not float(v).is_integer() or abs(v) >= 1e16 else str(int(v))


LammpsInputParser = LAMMPSInputParser


def parse_and_run(script_path, default_fp: Any = f64, arch: Any = None):
    """Parse and run a script; returns its Simulation."""
    p = LAMMPSInputParser(script_path, default_fp=default_fp, arch=arch)
    p.execute()
    return p.simulation
