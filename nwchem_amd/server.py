"""(T) serving with continuous batching (libtce_t.so's tce_t_server_*).

    srv = TriplesServer(slice_ms=50, max_active=16)
    sid = srv.add_system(tiles, blocks)        # host blocks stay referenced here
    t = srv.submit(sid, BF16X3)                # non-blocking
    r = srv.wait(t)                            # {'E_bt', 'E_pt', 'submit', 'admit', 'start', 'finish'}

One native worker thread owns the GPU; this module only passes pointers.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DP, LP, check
from .job import init
from .tiles import Blocks, TileModel


STATUS_TIMEOUT = -5  # TCE_T_ERR_STATE: the wait timed out


class TriplesServer:
    def __init__(self, device: int = 0, slice_ms: float = 50.0, max_active: int = 16,
                 reserve_bytes: int = 4 << 30):
        init(device)
        self._h = C.c_void_p()
        self._keep = []  # host blocks and tile descriptions of the systems
        check(_lib.lib().tce_t_server_create(device, slice_ms, max_active, reserve_bytes,
                                             C.byref(self._h)), "tce_t_server_create")

    def add_system(self, tiles: TileModel, b: Blocks) -> int:
        tc = tiles.c()
        self._keep.append((tiles, tc, b))
        sid = _lib.lib().tce_t_server_add_system(
            self._h, C.byref(tc), b.t1.ctypes.data_as(DP), b.t1h.ctypes.data_as(LP),
            b.t2.ctypes.data_as(DP), b.t2h.ctypes.data_as(LP), b.v2.ctypes.data_as(DP),
            b.v2h.ctypes.data_as(LP))
        check(min(sid, 0), "tce_t_server_add_system")
        return int(sid)

    def submit(self, system: int, precision: int = 0) -> int:
        t = _lib.lib().tce_t_server_submit(self._h, system, precision)
        check(int(min(t, 0)), "tce_t_server_submit")
        return int(t)

    def wait(self, ticket: int, timeout: float = -1.0) -> dict:
        e1, e2 = C.c_double(), C.c_double()
        tm = np.zeros(4)
        rc = _lib.lib().tce_t_server_wait(self._h, ticket, timeout, C.byref(e1), C.byref(e2),
                                          tm.ctypes.data_as(DP))
        if rc == STATUS_TIMEOUT and tm[3] == 0.0 and timeout >= 0:
            raise TimeoutError(f"request {ticket} not finished after {timeout} s")
        check(rc, "tce_t_server_wait")
        return {"E_bt": e1.value, "E_pt": e2.value, "submit": tm[0], "admit": tm[1],
                "start": tm[2], "finish": tm[3]}

    def stats(self) -> dict:
        sl, res = C.c_long(), C.c_int()
        check(_lib.lib().tce_t_server_stats(self._h, C.byref(sl), C.byref(res)), "tce_t_server_stats")
        return {"slices": sl.value, "resident": res.value}

    def close(self) -> None:
        if self._h:
            _lib.lib().tce_t_server_destroy(self._h)
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
