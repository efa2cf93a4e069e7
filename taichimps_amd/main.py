"""Command line: run a LAMMPS input script on the device DEM path
(reference: main.py).

    python -m taichimps_amd.main in.script [--fp f64|f32split]
    python -m taichimps_amd.main -in in.script
"""

from __future__ import annotations

import argparse
import sys

from .constants import f32split, f64
from .input import LAMMPSInputParser


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="taichimps_amd: granular DEM on MI355X")
    ap.add_argument("script", nargs="?", help="LAMMPS input script")
    ap.add_argument("-in", "--input", dest="in_script", help="LAMMPS-style -in <script>")
    ap.add_argument("--fp", "--dtype", dest="fp", default="f64", choices=["f64", "f32split"],
                    help="precision of the per-atom state")
    args = ap.parse_args(argv)
    script = args.script or args.in_script
    if not script:
        ap.print_help()
        return 0
    LAMMPSInputParser(script, default_fp=f64 if args.fp == "f64" else f32split).execute()
    return 0


if __name__ == "__main__":
    sys.exit(main())
