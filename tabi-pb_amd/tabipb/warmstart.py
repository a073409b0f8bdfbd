"""Learned initial guess for the GMRES solve: data generation, training and
evaluation around the GPU kernels (tabi_ws_*, tabi_ctx_ws_features).

The reference always starts DGMRES from zero (xvct = 0, src/main.f90:115).  Here
a per-face MLP predicts the tree-order solution [phi; dphi/dn] from local
features; GMRES then iterates to the same tolerance and falls back to the zero
start when the prediction's preconditioned residual is larger than that of zero.

Training data are solutions of the plain FP64 solver on seeded synthetic
molecules (``surface.synth_molecule`` + ``surface.ses_mesh``); each molecule
contributes a random subset of its faces.  Features and targets are scaled by
the molecule's charge scale s = 1/rms(b_1..b_n), so the model does not depend on
the overall charge magnitude.

Command line (on a machine with the GPU):
    python -m tabipb.warmstart train --nmol 1024 --out model.npz
    python -m tabipb.warmstart evaluate --model model.npz --nmol 128
"""
from __future__ import annotations

import argparse
import multiprocessing as mp
import os
import sys
import time

import numpy as np

from . import driver, hip, meshio, surface

NF, NH, NO = 12, 128, 2
FEATURES = ("s*b_phi", "s*b_dphi", "s*phi_5A", "s*dphi_5A", "s*phi_10A", "s*dphi_10A",
            "area", "curvature", "d_atom1", "d_atom2", "normal.dir_atom1", "burial_10A")
_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(_HERE, "data", "warmstart_mlp.npz")

# the physical set-up of the training molecules (as the benchmark's)
EPS_P, EPS_W, IONIC = 1.0, 80.0, 0.15
LEAF, ORDER, THETA = 500, 3, 0.8


def n_params() -> int:
    return 32 + NH * NF + NH + 2 * (NH * NH + NH) + NO * NH + NO


def init_params(mu, sig, ts, seed=0):
    """Packed parameters with LeCun-normal hidden weights, a small output layer
    and zero biases; mu/sig are the feature normalisation, ts the output scales."""
    rng = np.random.default_rng(seed)
    P = np.zeros(n_params(), dtype=np.float32)
    P[0:12] = mu
    P[12:24] = sig
    P[24:26] = ts
    o = 32
    for fan_out, fan_in, gain in ((NH, NF, 1.0), (NH, NH, 1.0), (NH, NH, 1.0), (NO, NH, 0.1)):
        w = rng.standard_normal((fan_out, fan_in)) * gain / np.sqrt(fan_in)
        P[o:o + fan_out * fan_in] = w.ravel()
        o += fan_out * fan_in + fan_out
    assert o == P.size
    return P


def save(path, params, meta=None):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=np.asarray(params, dtype=np.float32),
             meta=np.array(repr(meta or {})))


def load(path=DEFAULT_MODEL):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32)


def molecule(seed, natom, h=None):
    """One seeded synthetic molecule and its culled surface (CPU only)."""
    rng = np.random.default_rng(seed + 7919)
    if h is None:
        h = 0.5 if rng.random() < 0.5 else 0.6
    xyz, q, r = surface.synth_molecule(natom, seed=seed)
    v, nv, f = surface.ses_mesh(xyz, r, probe=1.4, h=h)
    f, _ = driver.cull_faces(v, f)
    fx, fn, fa = driver.face_geometry(v, nv, f)
    qp = xyz.astype(np.float32).astype(np.float64)
    qq = q.astype(np.float32).astype(np.float64)
    return dict(seed=seed, natom=natom, h=h, xyz=fx, nrm=fn, area=fa, qpos=qp, q=qq)


def _molecule_job(args):
    return molecule(*args)


def molecule_sizes(seeds, lo=300, hi=3000, salt=0):
    """Log-uniform atom counts in [lo, hi], one per seed."""
    out = []
    for s in seeds:
        u = np.random.default_rng(10007 * s + salt).random()
        out.append(int(round(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo))))))
    return out


def params_for(prec=0):
    usr = meshio.UsrData(maxparnode=LEAF, order=ORDER, mac=THETA)
    return hip.make_params(EPS_W / EPS_P, driver.kappa_of(IONIC, EPS_W), usr.mac, usr.order,
                           usr.maxparnode, prec)


def molecules(seeds, sizes, workers=8):
    """Generate molecules in worker processes (spawned: no GPU state is shared)."""
    jobs = list(zip(seeds, sizes))
    if workers <= 1:
        for j in jobs:
            yield molecule(*j)
        return
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        for m in pool.imap(_molecule_job, jobs, chunksize=1):
            yield m


def samples(m, nsample, rng, tol=1e-6):
    """Solve one molecule with the plain FP64 solver; return (features,
    targets, surface size, iterations) for a random subset of its faces."""
    S = hip.Surface(params_for(0), m["xyz"], m["nrm"], m["area"])
    try:
        b = S.rhs(m["qpos"], m["q"], EPS_P)
        x, it, err, ierr = S.gmres(b, tol=tol)
        feat, s = S.ws_features(m["qpos"], m["q"], EPS_P)
    finally:
        S.close()
    n = feat.shape[0]
    tgt = np.stack([x[:n] * s, x[n:] * s], axis=1).astype(np.float32)
    k = min(nsample, n)
    sel = rng.choice(n, size=k, replace=False)
    return feat[sel], tgt[sel], n, it


def build_dataset(seeds, sizes, nsample=8192, workers=8, log=print):
    rng = np.random.default_rng(12345)
    F, T = [], []
    t0 = time.time()
    for k, m in enumerate(molecules(seeds, sizes, workers)):
        f, t, n, it = samples(m, nsample, rng)
        F.append(f)
        T.append(t)
        if log and (k % 16 == 0 or k == len(seeds) - 1):
            log(f"[warmstart] data {k + 1}/{len(seeds)} seed={m['seed']} atoms={m['natom']} "
                f"faces={n} iters={it} ({time.time() - t0:.0f}s)")
    return np.concatenate(F), np.concatenate(T)


def fit(feat, tgt, epochs=40, batch=8192, lr=2e-3, seed=0, log=print):
    """Normalisation from the data, random init, Adam on the GPU."""
    mu = feat.mean(0).astype(np.float32)
    sig = feat.std(0).astype(np.float32)
    sig[sig < 1e-6] = 1.0
    ts = np.sqrt((tgt.astype(np.float64) ** 2).mean(0)).astype(np.float32)
    ts[ts < 1e-12] = 1.0
    P = init_params(mu, sig, ts, seed)
    M = hip.WsModel(P)
    try:
        losses = M.train(feat, tgt, epochs=epochs, batch=batch, lr=lr, seed=seed)
        P = M.params()
    finally:
        M.close()
    if log:
        log(f"[warmstart] trained {epochs} epochs on {feat.shape[0]} faces; loss "
            f"{losses[0]:.4g} -> {losses[-1]:.4g}")
    return P, losses


def evaluate(P, seeds, sizes, workers=8, tol=1e-4, prec=1, log=print):
    """Solve each molecule from zero and from the learned guess; returns per
    molecule (faces, iters zero, iters learned, guess used, dG zero, dG learned,
    dG FP64 reference, time zero ms, time learned ms)."""
    rows = []
    M = hip.WsModel(P)
    try:
        for m in molecules(seeds, sizes, workers):
            a = (m["xyz"], m["nrm"], m["area"], m["qpos"], m["q"], EPS_P)
            r64 = hip.solve(params_for(0), *a, tol=1e-6)
            r0 = hip.solve(params_for(prec), *a, tol=tol)
            r1 = hip.solve(params_for(prec), *a, tol=tol, model=M)
            rows.append((m["xyz"].shape[0], r0["iter"], r1["iter"], r1["ws_used"],
                         r0["soleng"], r1["soleng"], r64["soleng"],
                         r0["times_ms"]["total"], r1["times_ms"]["total"]))
            if log:
                log(f"[warmstart] eval seed={m['seed']} faces={rows[-1][0]} iters "
                    f"{r0['iter']} -> {r1['iter']} used={r1['ws_used']} dG rel err "
                    f"{abs(r1['soleng'] - r64['soleng']) / abs(r64['soleng']):.2e}")
    finally:
        M.close()
    return np.array(rows, dtype=np.float64)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--nmol", type=int, default=1024)
    t.add_argument("--first-seed", type=int, default=1000)
    t.add_argument("--nsample", type=int, default=8192)
    t.add_argument("--epochs", type=int, default=40)
    t.add_argument("--batch", type=int, default=8192)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--workers", type=int, default=8)
    t.add_argument("--min-atoms", type=int, default=300)
    t.add_argument("--max-atoms", type=int, default=3000)
    t.add_argument("--out", default=DEFAULT_MODEL)
    e = sub.add_parser("evaluate")
    e.add_argument("--model", default=DEFAULT_MODEL)
    e.add_argument("--nmol", type=int, default=128)
    e.add_argument("--first-seed", type=int, default=900000)
    e.add_argument("--workers", type=int, default=8)
    e.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    if a.cmd == "train":
        seeds = list(range(a.first_seed, a.first_seed + a.nmol))
        sizes = molecule_sizes(seeds, a.min_atoms, a.max_atoms)
        feat, tgt = build_dataset(seeds, sizes, a.nsample, a.workers, log)
        P, losses = fit(feat, tgt, a.epochs, a.batch, a.lr, log=log)
        save(a.out, P, dict(nmol=a.nmol, first_seed=a.first_seed, nsample=a.nsample,
                            epochs=a.epochs, batch=a.batch, lr=a.lr, min_atoms=a.min_atoms,
                            max_atoms=a.max_atoms, losses=[float(x) for x in losses]))
        log(f"[warmstart] model written to {a.out}")
    else:
        seeds = list(range(a.first_seed, a.first_seed + a.nmol))
        sizes = molecule_sizes(seeds, salt=1)
        rows = evaluate(load(a.model), seeds, sizes, a.workers, log=log)
        it0, it1 = rows[:, 1], rows[:, 2]
        err = np.abs(rows[:, 5] - rows[:, 6]) / np.abs(rows[:, 6])
        log(f"[warmstart] {len(rows)} held-out molecules: mean iters zero {it0.mean():.2f} "
            f"learned {it1.mean():.2f}; guess used {int(rows[:, 3].sum())}/{len(rows)}; "
            f"max dG rel err (learned, mixed) {err.max():.2e}; mean time zero "
            f"{rows[:, 7].mean():.1f} ms learned {rows[:, 8].mean():.1f} ms")
        if a.out:
            np.savetxt(a.out, rows, header="faces iters_zero iters_learned used dG_zero "
                       "dG_learned dG_fp64 ms_zero ms_learned")
    return 0


if __name__ == "__main__":
    sys.exit(main())
