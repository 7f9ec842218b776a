"""Error model of the reduced-precision (T) modes (test infrastructure).

The GPU's FP32 and split-BF16 modes differ from FP64 in two ways:

  representation  every d1/d2 operand is converted (FP32: fl32(x); BF16X3:
                  hi = bf16(fl32(x)), lo = bf16(fl32(x - hi)), and the
                  product hi*hi' + hi*lo' + lo*hi'; BF16X6: hi, mid, lo and
                  the six products hi*hi', hi*mid', mid*hi', hi*lo',
                  mid*mid', lo*hi').  This part is deterministic: the oracle
                  restates it exactly (oracle.arith("fp32" / "bf16x3" /
                  "bf16x6"), products exact in FP64, FP64 accumulation), and
                  the conversion itself is checked bit for bit against
                  oracle.split_bf16 / split3_bf16.
  accumulation    the products are summed in FP32 on the matrix cores.

Model of the accumulation error (probabilistic rounding-error analysis with
mean-independent rounding errors, Higham & Mary, SIAM J. Sci. Comput. 41
(2019) A2815): every element d_i of a t3 tile receives at most N_r rounding
errors (one per product and term: N_r = terms * (9 no + 9 nv), terms = 1 for
FP32, 3 for BF16X3, 6 for BF16X6), each of magnitude <= u |partial sum| <= u A_i, with
u = 2^-24 and A_i = sum of |products| of the element (oracle.arith("abs")).
The energy is E[T] = sum f d^2/D and E(T) = sum f d (d+s)/D, so to first order
dE = sum w_i dd_i with w_i = 2 f d_i / D_i (E[T]) or f (2 d_i + s_i)/D_i
(E(T)); the rounding errors are bounded, mean-independent steps of a
martingale, and the Azuma-Hoeffding inequality gives

    |E_gpu - E_emul| <= lam u sqrt(N_r sum_i w_i^2 A_i^2)
                        + sum_i |f/D_i| (lam sqrt(N_r) u A_i)^2

with probability >= 1 - 2 exp(-lam^2/2) (lam = 6: 3e-8); the second term
bounds the quadratic part of the energy.  E_emul is the oracle's energy in
the same representation, so the bound covers exactly what the oracle does
not restate.
"""
from __future__ import annotations

import numpy as np

import oracle

U32 = 2.0 ** -24
LAM = 6.0
TERMS = {"fp32": 1, "bf16x3": 3, "bf16x6": 6}
PRECISION_ID = {"fp32": 1, "bf16x3": 2, "bf16x6": 3}


def task_factor(restricted: bool, t) -> float:
    """ccsd_t_dot.F:51-65 (the oracle's task_factor)."""
    p4, p5, p6, h1, h2, h3 = t
    f = 2.0 if restricted else 1.0
    if p4 == p5 == p6:
        f /= 6.0
    elif p4 == p5 or p5 == p6:
        f /= 2.0
    if h1 == h2 == h3:
        f /= 6.0
    elif h1 == h2 or h2 == h3:
        f /= 2.0
    return f


def _denominators(tm, t):
    """D[p4,p5,p6,h1,h2,h3] = e_h1 + e_h2 + e_h3 - e_p4 - e_p5 - e_p6 of a task."""
    off = np.concatenate([[0], np.cumsum(tm.range)[:-1]])
    ev = [tm.evl[off[b - 1]:off[b - 1] + tm.range[b - 1]] for b in t]
    e4, e5, e6, e1, e2, e3 = ev
    return (-e4[:, None, None, None, None, None] - e5[None, :, None, None, None, None]
            - e6[None, None, :, None, None, None] + e1[None, None, None, :, None, None]
            + e2[None, None, None, None, :, None] + e3[None, None, None, None, None, :])


def task_sums(tm, t, s, d, a, precision: str, lam: float = LAM):
    """(sum N_r w^2 A^2 for E[T], the same for E(T), quadratic term) of one
    task from its FP64 tiles s, d and product magnitudes a (arrays
    [p4][p5][p6][h1][h2][h3]); evaluated one p4 slice at a time."""
    n_r = TERMS[precision] * 9 * (tm.no + tm.nv)
    f = task_factor(tm.restricted, t)
    D = _denominators(tm, t)
    l_bt = l_pt = q = 0.0
    for i in range(d.shape[0]):
        Di, di, si, ai = D[i], d[i], s[i], a[i]
        l_bt += float(np.sum((2.0 * f * di / Di * ai) ** 2))
        l_pt += float(np.sum((f * (2.0 * di + si) / Di * ai) ** 2))
        q += float(np.sum(np.abs(f / Di) * ai * ai))
    return n_r * l_bt, n_r * l_pt, q * (lam * U32) ** 2 * n_r


def bound_from_sums(l_bt, l_pt, q, lam: float = LAM):
    """The accumulation bound (E[T], E(T)) from task_sums (summed over tasks
    for a total)."""
    return np.array([lam * U32 * np.sqrt(l_bt) + q, lam * U32 * np.sqrt(l_pt) + q])


def accumulation_bounds(tm, b, ot, precision: str, tasks=None, lam: float = LAM):
    """Per-task accumulation bounds (n_tasks, 2) for (E[T], E(T)) and the bound
    of the totals (2,), from the oracle's FP64 tiles and product magnitudes."""
    if tasks is None:
        tasks = tm.tasks()
    per_task = np.zeros((len(tasks), 2))
    tot = np.zeros(3)
    for k, t in enumerate(tasks):
        s, d = oracle.task_tiles(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, t)
        with oracle.arith("abs"):
            _, a = oracle.task_tiles(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, t)
        sums = task_sums(tm, t, s, d, a, precision, lam)
        per_task[k] = bound_from_sums(*sums, lam=lam)
        tot += sums
    return per_task, bound_from_sums(*tot, lam=lam)


def element_bound(tm, a, precision: str, lam: float = LAM):
    """Per-element accumulation bound of a t3d tile: lam sqrt(N_r) u A_i."""
    n_r = TERMS[precision] * 9 * (tm.no + tm.nv)
    return lam * np.sqrt(n_r) * U32 * a


def emulated(tm, b, ot, mode: str):
    """(E[T], E(T), task energies) of the oracle in arithmetic `mode`."""
    with oracle.arith(mode):
        return oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
