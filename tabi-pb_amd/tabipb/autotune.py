"""Autotuner for the solve parameters the user is free to choose: leaf size
(maxparnode), Taylor order and the task granularity of the field kernels.

For one surface and charge set, every candidate is solved on the GPU and timed;
its solvation energy is compared with a converged reference solve: FP64, Taylor
order 6, MAC theta 0.5 and GMRES tolerance 1e-8 (the most accurate treecode the
configurations use; its matvec rows agree with the O(n^2) direct sum to ~1e-5
where p=3/theta=0.8 is off by ~5e-3), unless a reference energy is given.  The
error of the surface's own parameters against it is reported too.  The
fastest candidate whose energy stays within the error target wins.  The
reference method has no tuner (its parameters come from usrdata.in,
src/main.f90:45-47); the candidates are the values usrdata.in accepts.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

from . import hip

LEAVES = (100, 200, 300, 500, 700, 1000)
ORDERS = (2, 3, 4, 5, 6)
TASK_EXPONENTS = (0, 4, 8)       # flags bits 8-11: 2^16 (default), 2^14, 2^18 tasks


def timed_solve(prm, prob, tol, repeats):
    best, res = None, None
    for _ in range(repeats):
        t0 = time.perf_counter()
        res = hip.solve(prm, prob.xyz, prob.nrm, prob.area, prob.qpos, prob.q, prob.usr.epsp, tol)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best, res


def autotune(prob, leaves=LEAVES, orders=ORDERS, task_exponents=TASK_EXPONENTS, precision=1,
             err_target=1e-3, tol=1e-4, reference=None, repeats=2, log=None):
    """Returns dict(best=row, table=[rows], reference=dG, own_params_rel_err)
    where a row is dict(leaf, order, task_exp, ms, iters, dG, rel_err, ok) and
    own_params_rel_err is the FP64 solve at the surface's own parameters
    against the reference."""
    if reference is None:
        r = hip.solve(prob.params(0, order=6, theta=0.5), prob.xyz, prob.nrm, prob.area,
                      prob.qpos, prob.q, prob.usr.epsp, 1e-8)
        reference = r["soleng"]
    own = hip.solve(prob.params(0), prob.xyz, prob.nrm, prob.area, prob.qpos, prob.q,
                    prob.usr.epsp, tol)
    own_err = abs(own["soleng"] - reference) / abs(reference)
    rows = []
    # warm the library's allocator and code objects
    timed_solve(prob.params(precision), prob, tol, 1)
    for leaf, p, t in itertools.product(leaves, orders, task_exponents):
        prm = prob.params(precision, flags=t << 8, order=p, maxparnode=leaf)
        dt, res = timed_solve(prm, prob, tol, repeats)
        rel = abs(res["soleng"] - reference) / abs(reference)
        row = dict(leaf=leaf, order=p, task_exp=t, ms=dt * 1e3, iters=res["iter"],
                   dG=res["soleng"], rel_err=rel, ok=bool(res["ierr"] == 0 and rel <= err_target))
        rows.append(row)
        if log:
            log(f"[autotune] leaf={leaf} p={p} tasks=2^{10 + t if t else 16} {row['ms']:.1f} ms "
                f"iters={row['iters']} dG rel err {rel:.2e}{'' if row['ok'] else ' (rejected)'}")
    ok = [r for r in rows if r["ok"]]
    if not ok:
        raise RuntimeError("no candidate meets the error target")
    best = min(ok, key=lambda r: r["ms"])
    return dict(best=best, table=rows, reference=reference, err_target=err_target,
                precision=precision, reference_form="FP64, p=6, theta=0.5, tol 1e-8",
                own_params_rel_err=own_err)
