"""Writes the golden vectors of the (T) path from the reference's own code.

    python3 tests/golden/make_golden.py        (needs oracle/_ref built from
                                                the reference tree)

kernels.npz   the 27 kernels of ccsd_t_kernels_omp.F (sd_t_s1_1..9,
              sd_t_d1_1..9, sd_t_d2_1..9), compiled from the reference tree
              with strict FP (oracle/_ref/libsdt_ref.so), applied to the
              closed-form inputs of kernel_inputs() at two ragged shapes:
              the tile each one leaves behind.
energies.npz  E[T], E(T) and every task's pair for the systems in SYSTEMS,
              computed by the ccsd_t_gpu.F task loop (the oracle's dispatch)
              with the reference's compiled kernels and ccsd_t_dot as the
              arithmetic, plus SHA-256 digests of the input blocks and hash
              tables, so a test can tell an input change from a result change.

The inputs of the synthetic systems come from the library's Philox generator
(tce_t_synth_blocks); the spin-labelled ones from tests/spin_systems.py.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
ROOT = os.path.dirname(TESTS)
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

DP = C.POINTER(C.c_double)

# (h3d, h2d, h1d, p6d, p5d, p4d, K): every extent distinct where it can be,
# so a transposed operand or a swapped loop bound changes the result.
KERNEL_SHAPES = [(2, 3, 1, 4, 3, 2, 3), (3, 1, 2, 2, 5, 3, 2)]

# name -> (kind, parameters)
SYSTEMS = {
    "cfg1_no8_nv32": ("synthetic", dict(no=8, nv=32, to=8, tv=16, seed=1234)),
    "no7_nv13_to3_tv5": ("synthetic", dict(no=7, nv=13, to=3, tv=5, seed=11)),
    "no5_nv9_to2_tv4": ("synthetic", dict(no=5, nv=9, to=2, tv=4, seed=12)),
    "no10_nv22_to4_tv6": ("synthetic", dict(no=10, nv=22, to=4, tv=6, seed=13)),
    "no9_nv41_ragged": ("synthetic", dict(no=9, nv=41, to=0, tv=0, seed=15, ro=[6, 3], rv=[16, 16, 9])),
    "spin_sym_unrestricted": ("spin", dict(sizes_o=[3, 2], sizes_v=[4, 3], irreps_o=[0, 1],
                                           irreps_v=[1, 0], seed=5)),
    "spin_flip_unrestricted": ("spin", dict(sizes_o=[3, 2], sizes_v=[4, 3], seed=7,
                                            flip_symmetric=True)),
    "spin_flip_restricted": ("spin", dict(sizes_o=[3, 2], sizes_v=[4, 3], seed=7,
                                          flip_symmetric=True, restricted=True)),
}


def kernel_inputs(fam: str, shape, n: int):
    """Closed-form inputs (exact binary fractions, no RNG): the starting
    tile and the two operands of variant n of family fam."""
    h3d, h2d, h1d, p6d, p5d, p4d, K = shape
    n3 = h3d * h2d * h1d * p6d * p5d * p4d
    if fam == "s1":
        na, nb = p4d * h1d, h3d * h2d * p6d * p5d
    elif fam == "d1":
        na, nb = K * p4d * p5d * h1d, h3d * h2d * p6d * K
    else:
        na, nb = K * p4d * h1d * h2d, K * h3d * p6d * p5d
    salt = {"s1": 1, "d1": 2, "d2": 3}[fam] * 16 + n

    def seq(m, a, b):
        i = np.arange(m, dtype=np.int64)
        return ((i * a + b + 37 * salt) % 509 - 254) / 256.0

    return seq(n3, 97, 5), seq(na, 131, 11), seq(nb, 89, 23)


def build_system(kind: str, prm: dict):
    """(tiles for the oracle, blocks, tile model) of one golden system."""
    import oracle
    from nwchem_amd import tiles as tl

    if kind == "synthetic":
        tm = tl.TileModel.synthetic(prm["no"], prm["nv"], prm["to"], prm["tv"], prm["seed"],
                                    prm.get("ro"), prm.get("rv"))
        b = tl.synth_blocks(tm, prm["seed"])
        return oracle.Tiles(tm.range, tm.noab, tm.evl), b, tm
    import spin_systems

    s = spin_systems.SpinSystem(**prm)
    return s.oracle_tiles(), s.blocks, s.tiles


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main() -> int:
    import oracle

    R = oracle.ref_lib()
    out = {}
    for si, shape in enumerate(KERNEL_SHAPES):
        dims = [C.c_long(v) for v in shape]
        for fam in ("s1", "d1", "d2"):
            for n in range(1, 10):
                t3, a, b = kernel_inputs(fam, shape, n)
                f = getattr(R, f"sd_t_{fam}_{n}_")
                args = dims[:6] if fam == "s1" else dims
                f(*[C.byref(d) for d in args], t3.ctypes.data_as(DP), a.ctypes.data_as(DP),
                  b.ctypes.data_as(DP))
                out[f"shape{si}_{fam}_{n}"] = t3
    out["shapes"] = np.array(KERNEL_SHAPES, np.int64)
    np.savez_compressed(os.path.join(HERE, "kernels.npz"), **out)

    en, meta = {}, {}
    be = oracle.ref_backend()
    for name, (kind, prm) in SYSTEMS.items():
        ot, b, tm = build_system(kind, prm)
        e1, e2, te = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, backend=be)
        en[name + "_tasks"] = te
        en[name + "_total"] = np.array([e1, e2])
        meta[name] = {"kind": kind, "params": prm, "tasks": int(te.shape[0]),
                      "E_bt": float(e1).hex(), "E_pt": float(e2).hex(),
                      "inputs_sha256": digest(b.t1, b.t2, b.v2),
                      "tables_sha256": digest(b.t1h, b.t2h, b.v2h, tm.range, tm.evl)}
        print(f"{name}: {te.shape[0]} tasks, E[T]={e1:.15e} E(T)={e2:.15e}")
    np.savez_compressed(os.path.join(HERE, "energies.npz"), **en)
    with open(os.path.join(HERE, "energies.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
        f.write("\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
