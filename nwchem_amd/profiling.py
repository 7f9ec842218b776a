"""Profiling helpers (run under rocprofv3).

    python3 -m nwchem_amd.profiling copy [--mib 1024] [--repeats 1]

launches the library's calibration copy kernel (stream_copy): 8n bytes read
and 8n bytes written per launch, the yardstick that turns the FETCH_SIZE /
WRITE_SIZE counter readings of the (T) kernel into HBM bytes.
"""
from __future__ import annotations

import argparse
import sys

from . import _lib


def copy_probe(mib: int = 1024, repeats: int = 1) -> int:
    n = mib * 1024 * 1024 // 8
    L = _lib.lib()
    L.tce_t_copy_probe.argtypes = [_lib.C.c_long, _lib.C.c_int]
    rc = L.tce_t_copy_probe(n, repeats)
    _lib.check(rc, "tce_t_copy_probe")
    return n * 8


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["copy"])
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--repeats", type=int, default=1)
    a = ap.parse_args(argv)
    nbytes = copy_probe(a.mib, a.repeats)
    print(f"stream_copy: {a.repeats} launch(es), {nbytes} bytes read and written per launch")
    return 0


if __name__ == "__main__":
    sys.exit(main())
