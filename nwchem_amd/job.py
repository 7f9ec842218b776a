"""Whole-job (T) on the GPU: the host-side mirror of ccsd_t_gpu.F's surface.

    job = TriplesJob(tiles)                       # ccsd_t_gpu(d_t1,k_t1_offset,...)
    job.upload(blocks)                            # GA gets -> HBM, once
    e_bt, e_pt = job.run()                        # energy1 = E[T], energy2 = E(T)

Everything runs in libtce_t.so (HIP kernels on gfx950); this module only
passes pointers.  There is no CPU fallback: without a built library or a
GPU the calls raise.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DP, LP, check
from .tiles import Blocks, TileModel

FP64 = 0    # d1/d2 contractions on FP64 MFMA
FP32 = 1    # FP32 operands, FP32 MFMA and accumulation
BF16X3 = 2  # operands split into bf16 hi + lo, hi*hi + hi*lo + lo*hi, FP32 accumulation
BF16X6 = 3  # operands split into bf16 hi + mid + lo, the 6 products to 2^-24, FP32 accumulation

# Kernel-task tilings (tce_t_job_create_tiling)
TILING_REFERENCE = 0  # the reference tiles (the reference's work)
TILING_SUBTILE = 1    # box-sized sub-tiles of each reference tile (default)
TILING_ORBITAL = 2    # box-sized pieces of each spin/irrep run of orbitals; totals only

_initialised = False


def init(device: int = 0) -> None:
    global _initialised
    if not _initialised:
        check(_lib.lib().tce_t_init(device), "tce_t_init")
        _initialised = True


def device_memory() -> tuple[int, int]:
    """(free, total) bytes of the current device."""
    fr, tot = C.c_long(), C.c_long()
    check(_lib.lib().tce_t_device_memory(C.byref(fr), C.byref(tot)), "tce_t_device_memory")
    return fr.value, tot.value


class TriplesJob:
    def __init__(self, tiles: TileModel, device: int = 0, tiling: int | None = None):
        """tiling: TILING_REFERENCE, TILING_SUBTILE or TILING_ORBITAL (None:
        the library default, TILING_SUBTILE unless TCE_T_SUBTILE says
        otherwise).  TILING_ORBITAL jobs give E[T] and E(T) only."""
        init(device)
        self.tiles = tiles
        self._h = C.c_void_p()
        L = _lib.lib()
        if tiling is None:
            check(L.tce_t_job_create(C.byref(tiles.c()), C.byref(self._h)), "tce_t_job_create")
        else:
            check(L.tce_t_job_create_tiling(C.byref(tiles.c()), int(tiling), C.byref(self._h)),
                  "tce_t_job_create_tiling")
        self.ntasks = int(L.tce_t_job_num_tasks(self._h))
        self.tiling = int(L.tce_t_job_tiling(self._h))

    @property
    def device_bytes(self) -> int:
        return int(_lib.lib().tce_t_job_device_bytes(self._h))

    def upload(self, b: Blocks) -> None:
        L = _lib.lib()
        check(L.tce_t_job_upload(self._h, b.t1.ctypes.data_as(DP), b.t1h.ctypes.data_as(LP),
                                 b.t2.ctypes.data_as(DP), b.t2h.ctypes.data_as(LP),
                                 b.v2.ctypes.data_as(DP), b.v2h.ctypes.data_as(LP)),
              "tce_t_job_upload")

    def upload_lambda(self, y1: np.ndarray, y1h: np.ndarray, y2: np.ndarray, y2h: np.ndarray,
                      v2: np.ndarray, v2h: np.ndarray) -> None:
        """The Lambda amplitudes (tce_y1_offset / tce_y2_offset blocks) as
        the job's amplitudes, with the right-hand job's integrals: the
        left-hand job of lambda_ccsd_t."""
        arrs = [np.ascontiguousarray(a, dtype=dt) for a, dt in
                ((y1, np.float64), (y1h, np.int64), (y2, np.float64), (y2h, np.int64),
                 (v2, np.float64), (v2h, np.int64))]
        ptrs = [a.ctypes.data_as(DP if a.dtype == np.float64 else LP) for a in arrs]
        check(_lib.lib().tce_t_job_upload_lambda(self._h, *ptrs), "tce_t_job_upload_lambda")

    def upload_cr(self, t1: np.ndarray, t1h: np.ndarray, t2: np.ndarray, t2h: np.ndarray,
                  i1_1: np.ndarray, i1_1h: np.ndarray, i1_2: np.ndarray,
                  i1_2h: np.ndarray) -> None:
        """A CR-CCSD(T) job: t1, t2 and the two intermediates cr_ccsd_t_N
        builds before its task loop (i1(h11 p4 h1 h2) with its
        OFFSET_cr_ccsd_t_N_1_1 table, i1(p4 p5 h1 p12) with its
        OFFSET_cr_ccsd_t_N_2_1 table).  The job's t3 doubles are the moment
        tile, its singles the denominator tile of cr_ccsd_t_E."""
        arrs = [np.ascontiguousarray(a, dtype=dt) for a, dt in
                ((t1, np.float64), (t1h, np.int64), (t2, np.float64), (t2h, np.int64),
                 (i1_1, np.float64), (i1_1h, np.int64), (i1_2, np.float64), (i1_2h, np.int64))]
        ptrs = [a.ctypes.data_as(DP if a.dtype == np.float64 else LP) for a in arrs]
        check(_lib.lib().tce_t_job_upload_cr(self._h, *ptrs), "tce_t_job_upload_cr")

    def run(self, precision: int = FP64) -> tuple[float, float]:
        e1, e2 = C.c_double(0.0), C.c_double(0.0)
        check(_lib.lib().tce_t_job_run(self._h, precision, C.byref(e1), C.byref(e2)), "tce_t_job_run")
        return e1.value, e2.value

    def run_restart(self, state: dict | None = None, precision: int = FP64,
                    max_tiles: int = 0) -> dict:
        """ccsd_t_restart.F semantics: `state` = {'begin', 'table'} from a
        previous call (or None to start); returns the updated state plus
        't_energy' (E(T)) once every outer virtual tile is done."""
        nvab = int(self.tiles.range.size - self.tiles.noab)
        begin = C.c_long(1 if state is None else int(state["begin"]))
        table = np.zeros(nvab) if state is None else np.array(state["table"], dtype=float)
        te = C.c_double(0.0)
        check(_lib.lib().tce_t_job_run_restart(self._h, precision, C.byref(begin),
                                               table.ctypes.data_as(DP), max_tiles, C.byref(te)),
              "tce_t_job_run_restart")
        out = {"begin": begin.value, "table": table}
        if begin.value > nvab:
            out["t_energy"] = te.value
        return out

    def task_energies(self) -> np.ndarray:
        out = np.zeros((self.ntasks, 2))
        check(_lib.lib().tce_t_job_task_energies(self._h, out.ctypes.data_as(DP), self.ntasks),
              "tce_t_job_task_energies")
        return out

    def triples_tiles(self, h1b: int, h2b: int, h3b: int, p4b: int, p5b: int, p6b: int,
                      precision: int = FP64) -> tuple[np.ndarray, np.ndarray]:
        """(t3s, t3d) of the task with these 1-based tiles, as
        ccsd_t_singles_l / ccsd_t_doubles_l (toggle 2) leave them: arrays of
        shape (r(p4b), r(p5b), r(p6b), r(h1b), r(h2b), r(h3b)), i.e. the
        reference's triplesx(h3,h2,h1,p6,p5,p4) in Fortran order."""
        tiles = (p4b, p5b, p6b, h1b, h2b, h3b)
        if not all(1 <= t <= len(self.tiles.range) for t in tiles):
            # not tiles of the job: the library refuses the call (no buffers)
            check(_lib.lib().tce_t_job_triples_tiles(self._h, precision, h1b, h2b, h3b, p4b, p5b,
                                                     p6b, None, None), "tce_t_job_triples_tiles")
        shape = tuple(int(self.tiles.range[b - 1]) for b in tiles)
        s = np.zeros(shape)
        d = np.zeros(shape)
        check(_lib.lib().tce_t_job_triples_tiles(self._h, precision, h1b, h2b, h3b, p4b, p5b, p6b,
                                                 s.ctypes.data_as(DP), d.ctypes.data_as(DP)),
              "tce_t_job_triples_tiles")
        return s, d

    def triples_tiles_device(self, h1b: int, h2b: int, h3b: int, p4b: int, p5b: int, p6b: int,
                             precision: int = FP64) -> tuple[int, int, int]:
        """The same tiles left in HBM: (t3s device address, t3d device
        address, elements), valid until the job's next tile, Lambda or
        upload call."""
        s, d, n = C.c_void_p(), C.c_void_p(), C.c_long(0)
        check(_lib.lib().tce_t_job_triples_tiles_device(self._h, precision, h1b, h2b, h3b, p4b,
                                                        p5b, p6b, C.byref(s), C.byref(d),
                                                        C.byref(n)),
              "tce_t_job_triples_tiles_device")
        return s.value, d.value, n.value

    def triples_tiles_6dts(self, h1b: int, h2b: int, h3b: int, p4b: int, p5b: int, p6b: int,
                           istart: int, istop: int, jstart: int, jstop: int, maxp4: int,
                           maxp5: int, precision: int = FP64) -> tuple[np.ndarray, np.ndarray]:
        """The slice p4 in [istart, istop], p5 in [jstart, jstop] (1-based
        within the tiles) of the task's (t3s, t3d), as ccsd_t_singles_l_6dts /
        ccsd_t_doubles_l_6dts leave it: arrays of shape (maxp4, maxp5, r(p6b),
        r(h1b), r(h2b), r(h3b)) (triplesx(h3,h2,h1,p6,maxp5,maxp4), Fortran
        order), the slice in the leading corner."""
        r = self.tiles.range
        shape = (maxp4, maxp5) + tuple(int(r[b - 1]) for b in (p6b, h1b, h2b, h3b))
        s = np.zeros(shape)
        d = np.zeros(shape)
        check(_lib.lib().tce_t_job_triples_tiles_6dts(self._h, precision, h1b, h2b, h3b, p4b, p5b,
                                                      p6b, istart, istop, jstart, jstop, maxp4,
                                                      maxp5, s.ctypes.data_as(DP),
                                                      d.ctypes.data_as(DP)),
              "tce_t_job_triples_tiles_6dts")
        return s, d

    def work(self, precision: int = FP64) -> tuple[float, int]:
        """(flops, work items) of a run in `precision`: the reference's
        per-task flop count summed over the kernel tasks actually run (the
        box-sized sub-tiles of the reference tiles unless TCE_T_SUBTILE=0)."""
        f, n = C.c_double(0.0), C.c_long(0)
        check(_lib.lib().tce_t_job_work(self._h, precision, C.byref(f), C.byref(n)), "tce_t_job_work")
        return f.value, n.value

    def precision_bytes(self, precision: int) -> int:
        """Extra device bytes the first run in `precision` allocates."""
        return int(_lib.lib().tce_t_job_precision_bytes(self._h, precision))

    def operands(self, precision: int, tensor: int):
        """The reduced-precision copy of contraction tensor `tensor` (0 T2e,
        1 Vd2, 2 Wd1, 3 T2m) as the kernels read it: FP32 -> float32 array
        (rows, kout); BF16X3 -> (hi, lo), BF16X6 -> (hi, mid, lo) uint16 bf16
        patterns (rows, kout)."""
        L = _lib.lib()
        rows, kout = C.c_long(0), C.c_long(0)
        _lib.check(L.tce_t_job_operands(self._h, precision, tensor, None, None, None, C.byref(rows),
                                        C.byref(kout)), "tce_t_job_operands")
        shape = (rows.value, kout.value)
        if precision == FP32:
            a = np.zeros(shape, np.float32)
            _lib.check(L.tce_t_job_operands(self._h, precision, tensor, a.ctypes.data, None, None,
                                            None, None), "tce_t_job_operands")
            return a
        parts = [np.zeros(shape, np.uint16) for _ in range(2 if precision == BF16X3 else 3)]
        ptrs = [p.ctypes.data for p in parts] + [None] * (3 - len(parts))
        _lib.check(L.tce_t_job_operands(self._h, precision, tensor, *ptrs, None, None),
                   "tce_t_job_operands")
        return tuple(parts)

    @property
    def last_kernel_ms(self) -> float:
        return float(_lib.lib().tce_t_job_last_kernel_ms(self._h))

    @property
    def kernel(self) -> str:
        """Name of the contraction kernel the job runs."""
        k = int(_lib.lib().tce_t_job_kernel(self._h))
        return {0: "fused_triples", 1: "fused_triples_wg", 2: "fused_triples_mixed_fp64"}[k]

    @staticmethod
    def kernel_for(precision: int, fp64_kernel: str) -> str:
        """Kernel that runs the contractions in `precision`."""
        return fp64_kernel if precision == FP64 else \
            {FP32: "fused_triples_mixed_fp32", BF16X3: "fused_triples_mixed_bf16x3",
             BF16X6: "fused_triples_mixed_bf16x6"}[precision]

    def close(self) -> None:
        if self._h:
            _lib.lib().tce_t_job_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


LAMBDA_ORBITAL = 0          # t3 and y3 elements of the same orbitals
LAMBDA_REFERENCE_ORDER = 1  # element i of tdoubles with element i of the left tiles


def lambda_ccsd_t(right: TriplesJob, left: TriplesJob, f1: np.ndarray | None = None,
                  f1h: np.ndarray | None = None, pairing: int = LAMBDA_ORBITAL,
                  precision: int = FP64) -> tuple[float, float, np.ndarray]:
    """lambda_ccsd_t.F: (energy1, energy2, per-task energies) from the
    right-hand job (t1, t2, v2) and the left-hand job (upload_lambda).
    f1/f1h: the one-electron operator (tce_mo1e_offset blocks) or None to
    drop the Fock term."""
    e1, e2 = C.c_double(0.0), C.c_double(0.0)
    te = np.zeros((right.ntasks, 2))
    if (f1 is None) != (f1h is None):
        raise ValueError("f1 and f1h go together")
    fp = fh = None
    if f1 is not None:
        f1 = np.ascontiguousarray(f1, dtype=np.float64)
        f1h = np.ascontiguousarray(f1h, dtype=np.int64)
        fp, fh = f1.ctypes.data_as(DP), f1h.ctypes.data_as(LP)
    check(_lib.lib().tce_t_lambda_ccsd_t(right._h, left._h, fp, fh, int(pairing), int(precision),
                                         C.byref(e1), C.byref(e2), te.ctypes.data_as(DP),
                                         right.ntasks), "tce_t_lambda_ccsd_t")
    return e1.value, e2.value, te


def cr_ccsd_t(triples: TriplesJob, cr: TriplesJob,
              precision: int = FP64) -> tuple[float, float, dict]:
    """cr_ccsd_t.F: (energy1, energy2, terms) from a (T) job (upload) and a
    CR job of the same tiles (upload_cr); terms = num1, num2, den0, den1,
    den2, with energy1 = num1 / (1 + den1 + den0) and energy2 = num2 /
    (1 + den2 + den0).  The d1/d2 contractions run in `precision`."""
    e1, e2 = C.c_double(0.0), C.c_double(0.0)
    terms = np.zeros(5)
    check(_lib.lib().tce_t_cr_ccsd_t(triples._h, cr._h, int(precision), C.byref(e1),
                                     C.byref(e2), terms.ctypes.data_as(DP)), "tce_t_cr_ccsd_t")
    return e1.value, e2.value, dict(zip(("num1", "num2", "den0", "den1", "den2"),
                                        (float(x) for x in terms)))


FAMILIES = {0: "fused_triples", 1: "fused_triples_wg", 2: "fused_triples_mixed",
            3: "fused_triples_mixed_fp64"}


class TriplesBatch:
    """Several uploaded TriplesJobs run by one persistent launch per call
    (autotuned on the first call in a precision, then replayed as a
    hipGraph).  The jobs must stay open while the batch exists."""

    def __init__(self, jobs):
        self.jobs = list(jobs)
        arr = (C.c_void_p * len(self.jobs))(*[j._h for j in self.jobs])
        self._h = C.c_void_p()
        check(_lib.lib().tce_t_batch_create(arr, len(self.jobs), C.byref(self._h)), "tce_t_batch_create")

    def run(self, precision: int = FP64) -> np.ndarray:
        """(njobs, 2) array of E[T], E(T)."""
        out = np.zeros((len(self.jobs), 2))
        check(_lib.lib().tce_t_batch_run(self._h, precision, out.ctypes.data_as(DP)), "tce_t_batch_run")
        return out

    @property
    def last_ms(self) -> float:
        return float(_lib.lib().tce_t_batch_last_ms(self._h))

    def plan(self, precision: int) -> dict:
        fam, co = C.c_int(), C.c_int()
        tms = np.zeros(6)
        check(_lib.lib().tce_t_batch_plan(self._h, precision, C.byref(fam), C.byref(co),
                                          tms.ctypes.data_as(DP)), "tce_t_batch_plan")
        sl, sms = C.c_int(), np.zeros(3)
        check(_lib.lib().tce_t_batch_plan_schedule(self._h, precision, C.byref(sl),
                                                   sms.ctypes.data_as(DP)),
              "tce_t_batch_plan_schedule")
        return {"kernel": FAMILIES[fam.value], "cost_order": bool(co.value),
                "tuned_ms": [float(x) for x in tms if x >= 0], "slice": int(sl.value),
                "slice_ms": [float(x) for x in sms if x >= 0]}

    def close(self) -> None:
        if self._h:
            _lib.lib().tce_t_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
