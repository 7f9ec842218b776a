"""ctypes binding of libtce_t.so (the HIP library built from nwchem_amd/csrc).

The library is loaded from the package directory only; there is no CPU
fallback.  If it has not been built, importing the GPU entry points raises.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtce_t.so")
CSRC = os.path.join(_HERE, "csrc")

_lib = None

LP = C.POINTER(C.c_long)
DP = C.POINTER(C.c_double)


class TilesC(C.Structure):
    """Mirror of tce_t_tiles (include/tce_t.h)."""

    _fields_ = [
        ("noab", C.c_long),
        ("nvab", C.c_long),
        ("spin", LP),
        ("sym", LP),
        ("range", LP),
        ("offset", LP),
        ("alpha", LP),
        ("evl", DP),
        ("restricted", C.c_int),
    ]


class IntorbC(C.Structure):
    """Mirror of tce_t_intorb (include/tce_t.h)."""
    _fields_ = [("noa", C.c_long), ("nva", C.c_long), ("b2am", LP), ("spin_alpha", LP),
                ("sym_alpha", LP), ("range_alpha", LP), ("irrep_v", C.c_long)]


def build(force: bool = False) -> str:
    """Compile libtce_t.so in-tree (hipcc, gfx950)."""
    cmd = ["make", "-C", CSRC]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    """Load libtce_t.so, building it only when nothing is built yet."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    _declare(L)
    _lib = L
    return L


def _declare(L: C.CDLL) -> None:
    L.tce_t_last_error.restype = C.c_char_p
    L.tce_t_init.argtypes = [C.c_int]
    L.tce_t_job_create.argtypes = [C.POINTER(TilesC), C.POINTER(C.c_void_p)]
    L.tce_t_job_num_tasks.argtypes = [C.c_void_p]
    L.tce_t_job_num_tasks.restype = C.c_long
    L.tce_t_job_device_bytes.argtypes = [C.c_void_p]
    L.tce_t_job_device_bytes.restype = C.c_long
    L.tce_t_job_upload.argtypes = [C.c_void_p, DP, LP, DP, LP, DP, LP]
    L.tce_t_job_run.argtypes = [C.c_void_p, C.c_int, DP, DP]
    L.tce_t_job_task_energies.argtypes = [C.c_void_p, DP, C.c_long]
    L.tce_t_job_last_kernel_ms.argtypes = [C.c_void_p]
    L.tce_t_job_last_kernel_ms.restype = C.c_double
    L.tce_t_job_precision_bytes.argtypes = [C.c_void_p, C.c_int]
    L.tce_t_job_precision_bytes.restype = C.c_long
    L.tce_t_job_operands.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long)]
    L.tce_t_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p)]
    L.tce_t_batch_run.argtypes = [C.c_void_p, C.c_int, DP]
    L.tce_t_batch_last_ms.argtypes = [C.c_void_p]
    L.tce_t_batch_last_ms.restype = C.c_double
    L.tce_t_batch_plan.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), DP]
    L.tce_t_batch_plan_schedule.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), DP]
    L.tce_t_batch_destroy.argtypes = [C.c_void_p]
    L.tce_t_server_create.argtypes = [C.c_int, C.c_double, C.c_int, C.c_long, C.POINTER(C.c_void_p)]
    L.tce_t_server_add_system.argtypes = [C.c_void_p, C.POINTER(TilesC), DP, LP, DP, LP, DP, LP]
    L.tce_t_server_submit.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.tce_t_server_submit.restype = C.c_long
    L.tce_t_server_wait.argtypes = [C.c_void_p, C.c_long, C.c_double, DP, DP, DP]
    L.tce_t_server_stats.argtypes = [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_int)]
    L.tce_t_server_destroy.argtypes = [C.c_void_p]
    L.tce_t_job_run_restart.argtypes = [C.c_void_p, C.c_int, LP, DP, C.c_long, DP]
    L.tce_t_job_triples_tiles.argtypes = [C.c_void_p, C.c_int] + [C.c_long] * 6 + [DP, DP]
    L.tce_t_job_triples_tiles_6dts.argtypes = [C.c_void_p, C.c_int] + [C.c_long] * 12 + [DP, DP]
    L.tce_t_job_triples_tiles_device.argtypes = [C.c_void_p, C.c_int] + [C.c_long] * 6 + [
        C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), LP]
    L.tce_t_job_upload_lambda.argtypes = [C.c_void_p, DP, LP, DP, LP, DP, LP]
    L.tce_t_job_upload_cr.argtypes = [C.c_void_p, DP, LP, DP, LP, DP, LP, DP, LP]
    L.tce_t_cr_ccsd_t.argtypes = [C.c_void_p, C.c_void_p, C.c_int, DP, DP, DP]
    L.tce_t_lambda_ccsd_t.argtypes = [C.c_void_p, C.c_void_p, DP, LP, C.c_int, C.c_int, DP, DP,
                                      DP, C.c_long]
    L.tce_t_device_memory.argtypes = [C.POINTER(C.c_long), C.POINTER(C.c_long)]
    L.tce_t_job_kernel.argtypes = [C.c_void_p]
    L.tce_t_job_kernel.restype = C.c_int
    L.tce_t_job_work.argtypes = [C.c_void_p, C.c_int, DP, C.POINTER(C.c_long)]
    L.tce_t_job_create_tiling.argtypes = [C.POINTER(TilesC), C.c_int, C.POINTER(C.c_void_p)]
    L.tce_t_job_tiling.argtypes = [C.c_void_p]
    L.tce_t_job_destroy.argtypes = [C.c_void_p]
    L.tce_t_pack_dense.argtypes = [C.POINTER(TilesC), DP, LP, DP, LP, DP, LP] + [DP] * 6
    L.tce_t_tile_ranges.argtypes = [C.c_long, C.c_long, LP]
    L.tce_t_tile_ranges.restype = C.c_long
    for f in ("tce_t_t1_offset", "tce_t_t2_offset", "tce_t_v2_offset", "tce_t_y1_offset",
              "tce_t_y2_offset", "tce_t_f1_offset"):
        getattr(L, f).argtypes = [C.POINTER(TilesC), LP, LP]
        getattr(L, f).restype = C.c_long
    L.tce_t_hash.argtypes = [LP, C.c_long]
    L.tce_t_hash.restype = C.c_long
    L.tce_t_v2_alpha_offset.argtypes = [C.POINTER(IntorbC), C.c_long, LP, LP]
    L.tce_t_v2_alpha_offset.restype = C.c_long
    L.tce_t_v2_from_intorb.argtypes = [C.POINTER(TilesC), C.POINTER(IntorbC), DP, LP, LP, DP]
    L.tce_t_synth_evl.argtypes = [C.c_ulonglong, C.c_long, C.c_long, DP]
    L.tce_t_synth_blocks.argtypes = [C.POINTER(TilesC), C.c_ulonglong, DP, LP, DP, LP, DP, LP]
    L.tce_t_synth_dense.argtypes = [C.c_ulonglong, C.c_long, C.c_long, DP, DP, DP, DP, DP]
    L.tce_t_philox4x32.argtypes = [C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.POINTER(C.c_uint)]
    L.tce_t_philox4x32.restype = None
    L.tce_t_legacy_status.restype = C.c_int
    L.tce_t_legacy_backend.argtypes = [C.c_int]
    L.tce_t_legacy_backend.restype = C.c_int
    LPtr = C.POINTER(C.c_long)
    for fam, nd in (("s1", 6), ("d1", 7), ("d2", 7)):
        for n in range(1, 10):
            fn = getattr(L, f"sd_t_{fam}_{n}_cuda_")
            fn.argtypes = [LPtr] * nd + [DP, DP, DP]
            fn.restype = None
    for fam, nd in (("s1", 6), ("d1", 7), ("d2", 7)):  # the CPU surface
        for n in range(1, 10):
            fn = getattr(L, f"sd_t_{fam}_{n}_")
            fn.argtypes = [LPtr] * nd + [DP, DP, DP]
            fn.restype = None
    L.compute_en_.argtypes = [DP, DP] + [DP] * 6 + [LPtr] * 6 + [DP, DP]
    L.compute_en_.restype = None
    for f in ("dev_mem_d_", "dev_mem_s_"):
        getattr(L, f).argtypes = [LPtr] * 6
        getattr(L, f).restype = None
    for f in ("dev_release_", "initmemmodule_", "finalizememmodule_"):
        getattr(L, f).argtypes = []
        getattr(L, f).restype = None
    L.check_device_.argtypes = [LPtr]
    L.check_device_.restype = C.c_int
    L.device_init_.argtypes = [LPtr, LPtr]
    L.device_init_.restype = C.c_int


def check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().tce_t_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed with status {status}: {msg}")


def exported_symbols() -> list[str]:
    """Names the shared object exports (nm -D)."""
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    return [ln.split()[-1] for ln in out.splitlines() if ln.strip()]
