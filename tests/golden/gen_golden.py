"""Regenerate the golden fixtures from the reference program.

Runs oracle/_ref/ref_harness.exe -- the reference's own Fortran compiled from
its sources by oracle/Makefile -- on each case of cases.py and stores what it
computes as compressed .npz files next to this script:

  orderarr (1-based), leaf ranges, tree-order right-hand side, y = A x and
  z = M^-1 x for the harness test vector, GMRES iter/err/ierr, the solution,
  the solvation/Coulomb energies, and the vertex values of the
  surface_potential.dat that the reference's output_potential writes (and, up
  to 8,000 faces, the file's text).

Vectors are stored in full up to 8,000 faces, every 8th entry up to 50,000
faces and every 64th entry beyond (plus norms); beyond 50,000 faces the
permutation is stored as a SHA-256 of its int32 bytes plus every 64th entry.
Usage: python tests/golden/gen_golden.py [case ...]   (default: the small cases)
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import cases  # noqa: E402
from tabipb import meshio  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness.exe")


def read_outputs(d):
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
    out = dict(n=n, kappa=kappa, eps=eps, txyz=txyz, tnrm=tq, tarea=tarea, orderarr=order,
               bvct=bv, leaf_beg=lb, leaf_end=le, x=x, y=y)
    mp = os.path.join(d, "ms.bin")
    if os.path.exists(mp):
        ms = open(mp, "rb").read()
        out["z"] = np.frombuffer(ms[4 + 8 * n2:], np.float64).copy()
    sp = os.path.join(d, "solve.bin")
    if os.path.exists(sp):
        s = open(sp, "rb").read()
        it, ierr = np.frombuffer(s[:8], np.int32)
        err, sol, cou = np.frombuffer(s[8:32], np.float64)
        out.update(iter=int(it), ierr=int(ierr), err=float(err), soleng=float(sol),
                   couleng=float(cou), xvct=np.frombuffer(s[32:], np.float64).copy())
    pp = os.path.join(d, "surface_potential.dat")
    if os.path.exists(pp):
        vpos, vnrm, vpot, faces = meshio.read_surface_potential(pp)
        out.update(vpot=vpot, pot_faces=faces)
    return out


def run_case(name):
    v, nv, f, atoms, usr = cases.build(name)
    opt = cases.harness_opts(name)
    args = [f"tol={opt['tol']!r}", f"nrmax={opt['nrmax']}", "time"]
    if not opt["solve"]:
        args.append("nosolve")
    else:
        args.append("potential")
    if not opt["msolve"]:
        args.append("nomsolve")
    with tempfile.TemporaryDirectory() as d:
        cases.stage(name, d)
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {HARNESS} {' '.join(args)}"],
                           cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stdout + r.stderr)
        o = read_outputs(d)
        times = [ln for ln in r.stdout.splitlines() if ln.startswith("harness-time")]
        pp = os.path.join(d, "surface_potential.dat")
        pot_text = open(pp).read() if os.path.exists(pp) else None
    n = o["n"]
    stride = 1 if n <= 8000 else (8 if n <= 50000 else 64)
    s = slice(None, None, stride)
    rec = dict(
        input_sha256=np.array(cases.input_hash(v, nv, f, atoms)),
        n=np.int64(n), kappa=np.float64(o["kappa"]), eps=np.float64(o["eps"]),
        leaf_beg=o["leaf_beg"], leaf_end=o["leaf_end"],
        bvct=o["bvct"][s], x=o["x"][s], y=o["y"][s],
        y_norm=np.float64(np.linalg.norm(o["y"])), stride=np.int64(stride),
        tol=np.float64(opt["tol"]), nrmax=np.int64(opt["nrmax"]),
        harness_times=np.array("\n".join(times)))
    order = np.ascontiguousarray(o["orderarr"], np.int32)
    if n <= 50000:
        rec["orderarr"] = order
    else:
        rec["orderarr_sha256"] = np.array(hashlib.sha256(order.tobytes()).hexdigest())
        rec["orderarr_sample"] = order[::64]
    if "z" in o:
        rec.update(z=o["z"][s], z_norm=np.float64(np.linalg.norm(o["z"])))
    if "iter" in o:
        rec.update(iter=np.int64(o["iter"]), ierr=np.int64(o["ierr"]), err=np.float64(o["err"]),
                   soleng=np.float64(o["soleng"]), couleng=np.float64(o["couleng"]),
                   xvct=o["xvct"][s], xvct_norm=np.float64(np.linalg.norm(o["xvct"])))
    if "vpot" in o:
        rec.update(vpot=o["vpot"][s], vpot_nan=np.flatnonzero(np.isnan(o["vpot"][:, 0])),
                   pot_nface=np.int64(len(o["pot_faces"])))
    if stride == 1:
        rec.update(txyz=o["txyz"], tnrm=o["tnrm"], tarea=o["tarea"])
        if pot_text is not None:
            rec["pot_text"] = np.array(pot_text)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **rec)
    print(f"{name}: n={n} leaves={len(o['leaf_beg'])} iter={o.get('iter')} "
          f"ierr={o.get('ierr')} soleng={o.get('soleng')}", flush=True)
    for t in times:
        print("   ", t, flush=True)


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        sys.exit("build the reference first: make -C oracle ref")
    for name in (sys.argv[1:] or cases.SMALL):
        run_case(name)
