"""Export the library for use as the Weakrates2 thorn of Cactus/THC
(the role of the reference's utils/export_thc.py, README.md:54-67).

The reference exports its header-only sources with macro substitutions
(KOKKOS_INLINE_FUNCTION -> inline, BS_PRINTF -> printf).  Here the headers
have no Kokkos or printf macros: the thorn's src/ receives the C ABI header,
the header-only C++ API with the reference's names, and the built library,
and the thorn links against it.

    python -m bns_nurates_amd.export_thc /path/to/Weakrates2/src

Files already at the destination are replaced only when they differ; the
script prints what changed and the flags to add to the thorn's build.
"""
from __future__ import annotations

import argparse
import filecmp
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FILES = [os.path.join(ROOT, "include", "bns_nurates_amd.h"),
         os.path.join(ROOT, "include", "bns_nurates_amd.hpp"),
         os.path.join(HERE, "libbnsamd.so")]


def export(destination: str, log=print) -> list[str]:
    """Copy the headers and the library into destination; returns the
    names of the files written."""
    if not os.path.isdir(destination):
        raise FileNotFoundError(f"destination {destination} is not a directory")
    missing = [f for f in FILES if not os.path.exists(f)]
    if missing:
        raise FileNotFoundError(f"not built: {missing} (run __graft_entry__.build())")
    written = []
    for src in FILES:
        dst = os.path.join(destination, os.path.basename(src))
        if os.path.exists(dst) and filecmp.cmp(src, dst, shallow=False):
            continue
        shutil.copy2(src, dst)
        written.append(os.path.basename(src))
        log(f"exported {os.path.basename(src)}")
    log(f"thorn build flags: -I{destination} -L{destination} -lbnsamd -Wl,-rpath,{destination}")
    return written


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("destination", help="the thorn's src directory")
    args = ap.parse_args(argv)
    try:
        export(args.destination)
    except FileNotFoundError as e:
        print(f"export_thc: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
