"""The C ABI without a GPU: the library loads, exports exactly what
include/tmps.h declares, and the ctypes mirrors of its structs have the C
layout (sizes and offsets printed by a C compiler from the header itself).
No entry point that touches the device is called here.
"""

from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tmps.h")
LIB = os.path.join(ROOT, "taichimps_amd", "libtmps_hip.so")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tmps_[a-z0-9_]+)\s*\(", text)))


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True,
                         capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


@pytest.fixture(scope="module")
def lib_path():
    if not os.path.exists(LIB):
        pytest.fail("libtmps_hip.so is not built; run __graft_entry__.build() first")
    return LIB


def test_header_declares_the_expected_entry_points():
    fns = header_functions()
    assert len(fns) >= 40
    for must in ("tmps_create", "tmps_neigh_build", "tmps_pair_compute", "tmps_nve_initial",
                 "tmps_virial_reduce", "tmps_servo_end_of_step", "tmps_run",
                 "tmps_adjoint_forward", "tmps_adjoint_backward", "tmps_batch_create",
                 "tmps_batch_run", "tmps_batch_tuning", "tmps_batch_destroy",
                 "tmps_contact_atom", "tmps_fabric", "tmps_stress_atom", "tmps_history_seed",
                 "tmps_wall_create", "tmps_wall_post_force"):
        assert must in fns


def test_library_exports_every_declared_symbol(lib_path):
    exported = exported_symbols(lib_path)
    missing = [f for f in header_functions() if f not in exported]
    assert not missing, f"declared in tmps.h but not exported: {missing}"
    extra = sorted(s for s in exported if s.startswith("tmps_") and s not in header_functions())
    assert not extra, f"exported but not declared in tmps.h: {extra}"


def test_python_binding_covers_the_header():
    from taichimps_amd import _native
    assert sorted(_native.SYMBOLS) == header_functions()


def test_library_loads_and_reports_version(lib_path):
    from taichimps_amd import _native
    L = _native.load()
    assert b"gfx950" in L.tmps_version()
    for name in header_functions():
        assert hasattr(L, name)


_LAYOUT_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "tmps.h"
#define F(T, m) printf(#T " " #m " %zu\n", offsetof(T, m))
int main(void) {
    printf("tmps_pair_params sizeof %zu\n", sizeof(tmps_pair_params));
    F(tmps_pair_params, kn); F(tmps_pair_params, xmu); F(tmps_pair_params, normal_model);
    F(tmps_pair_params, g_kn); F(tmps_pair_params, g_mu);
    printf("tmps_servo_params sizeof %zu\n", sizeof(tmps_servo_params));
    F(tmps_servo_params, max_rate); F(tmps_servo_params, style); F(tmps_servo_params, rate);
    F(tmps_servo_params, amp); F(tmps_servo_params, pgain);
    printf("tmps_counters sizeof %zu\n", sizeof(tmps_counters));
    F(tmps_counters, npairs); F(tmps_counters, servo_nsteps);
    printf("tmps_wall_params sizeof %zu\n", sizeof(tmps_wall_params));
    F(tmps_wall_params, shear_axis); F(tmps_wall_params, kn); F(tmps_wall_params, xmu);
    F(tmps_wall_params, geometry); F(tmps_wall_params, has_bounds); F(tmps_wall_params, c1);
    F(tmps_wall_params, axis_hi);
    return 0;
}
"""


def test_struct_layouts_match_the_header(tmp_path):
    from taichimps_amd import _native
    src = tmp_path / "probe.c"
    exe = tmp_path / "probe"
    src.write_text(_LAYOUT_PROBE)
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    mirror = {"tmps_pair_params": _native.PairParams, "tmps_servo_params": _native.ServoParams,
              "tmps_counters": _native.Counters, "tmps_wall_params": _native.WallParams}
    for line in lines.splitlines():
        struct, member, value = line.split()
        py = mirror[struct]
        got = C.sizeof(py) if member == "sizeof" else getattr(py, member).offset
        assert got == int(value), f"{struct}.{member}: ctypes {got}, C {value}"


def test_oracle_structs_match_the_library_structs():
    """The oracle's ctypes structs feed the same parameter sets to both sides."""
    from oracle import oracle as O
    from taichimps_amd import _native
    assert [f[0] for f in O.PairParams._fields_] == [f[0] for f in _native.PairParams._fields_]
    assert [f[0] for f in O.ServoParams._fields_] == [f[0] for f in _native.ServoParams._fields_]
    assert [f[0] for f in O.WallParams._fields_] == [f[0] for f in _native.WallParams._fields_]
