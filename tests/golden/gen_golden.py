"""Regenerate the golden fixtures from the reference program.

Runs oracle/_ref/ref_harness.exe -- the reference's own Fortran compiled from
its sources by oracle/Makefile -- on each case of cases.py and stores what it
computes as compressed .npz files next to this script:

  orderarr (1-based), leaf ranges, tree-order right-hand side, y = A x and
  z = M^-1 x for the harness test vector, GMRES iter/err/ierr, the solution and
  the solvation/Coulomb energies.

Large vectors are stored in full for small cases and as every 8th entry (plus
norms) for the 20,480-face sphere.  Usage: python tests/golden/gen_golden.py [case ...]
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import cases  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness.exe")


def read_outputs(d, n_expect=None):
    g = open(os.path.join(d, "geom.bin"), "rb").read()
    n = int(np.frombuffer(g[:4], np.int32)[0])
    kappa, eps = np.frombuffer(g[4:20], np.float64)
    off = 20
    def take(dt, cnt):
        nonlocal off
        a = np.frombuffer(g[off:off + cnt * np.dtype(dt).itemsize], dt)
        off += cnt * np.dtype(dt).itemsize
        return a.copy()
    txyz = take(np.float64, 3 * n).reshape(n, 3)
    tq = take(np.float64, 3 * n).reshape(n, 3)
    tarea = take(np.float64, n)
    order = take(np.int32, n)
    bv = take(np.float64, 2 * n)
    lv = open(os.path.join(d, "leaves.bin"), "rb").read()
    nl = int(np.frombuffer(lv[:4], np.int32)[0])
    lb = np.frombuffer(lv[4:4 + 4 * nl], np.int32).copy()
    le = np.frombuffer(lv[4 + 4 * nl:4 + 8 * nl], np.int32).copy()
    mv = open(os.path.join(d, "mv.bin"), "rb").read()
    n2 = int(np.frombuffer(mv[:4], np.int32)[0])
    x = np.frombuffer(mv[4:4 + 8 * n2], np.float64).copy()
    y = np.frombuffer(mv[4 + 8 * n2:], np.float64).copy()
    ms = open(os.path.join(d, "ms.bin"), "rb").read()
    z = np.frombuffer(ms[4 + 8 * n2:], np.float64).copy()
    out = dict(n=n, kappa=kappa, eps=eps, txyz=txyz, tnrm=tq, tarea=tarea, orderarr=order,
               bvct=bv, leaf_beg=lb, leaf_end=le, x=x, y=y, z=z)
    sp = os.path.join(d, "solve.bin")
    if os.path.exists(sp):
        s = open(sp, "rb").read()
        it, ierr = np.frombuffer(s[:8], np.int32)
        err, sol, cou = np.frombuffer(s[8:32], np.float64)
        out.update(iter=int(it), ierr=int(ierr), err=float(err), soleng=float(sol),
                   couleng=float(cou), xvct=np.frombuffer(s[32:], np.float64).copy())
    return out


def run_case(name):
    v, nv, f, atoms, usr = cases.build(name)
    with tempfile.TemporaryDirectory() as d:
        cases.stage(name, d)
        env = dict(os.environ)
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {HARNESS}"], cwd=d,
                           capture_output=True, text=True, env=env)
        if r.returncode != 0:
            raise RuntimeError(r.stdout + r.stderr)
        o = read_outputs(d)
    n = o["n"]
    big = n > 8000
    s = slice(None, None, 8) if big else slice(None)
    rec = dict(
        input_sha256=np.array(cases.input_hash(v, nv, f, atoms)),
        n=np.int64(n), kappa=np.float64(o["kappa"]), eps=np.float64(o["eps"]),
        orderarr=o["orderarr"], leaf_beg=o["leaf_beg"], leaf_end=o["leaf_end"],
        bvct=o["bvct"][s], x=o["x"][s], y=o["y"][s], z=o["z"][s],
        y_norm=np.float64(np.linalg.norm(o["y"])), z_norm=np.float64(np.linalg.norm(o["z"])),
        stride=np.int64(8 if big else 1),
        iter=np.int64(o["iter"]), ierr=np.int64(o["ierr"]), err=np.float64(o["err"]),
        soleng=np.float64(o["soleng"]), couleng=np.float64(o["couleng"]),
        xvct=o["xvct"][s])
    if not big:
        rec.update(txyz=o["txyz"], tnrm=o["tnrm"], tarea=o["tarea"])
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **rec)
    print(f"{name}: n={n} leaves={len(o['leaf_beg'])} iter={o['iter']} "
          f"soleng={o['soleng']:.12f}")


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        sys.exit("build the reference first: make -C oracle ref")
    for name in (sys.argv[1:] or list(cases.CASES)):
        run_case(name)
