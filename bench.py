"""Benchmark: PB solves/hour on one MI355X.

A step is one complete PB solve through libtabi_hip.so (tabi_solve_dev, with the
surface and charges already in device memory): tree build, interaction lists,
leaf-block LU factorisation, right-hand side, GMRES to
tol 1e-4 (with the reference's MAXL=10/NRMAX=10 left preconditioning and
stopping rule), and the solvation energy.  Three forms are timed: FP64 from a
zero start, mixed precision from a zero start, and mixed precision from the
learned initial guess (features + MLP inference inside the timed solve); the
reported value is the fastest form whose energy is within 1e-3 of the FP64
solve's.

Besides the JSON line's throughput, the run times the configuration's other
expansion (p = 6, theta = 0.5) in both precisions, measures the dominant
per-matvec kernel live with HIP events on the library's stream
(tabi_ctx_kernel_times) and reports its algorithmic FLOP rate against the
vector-ALU peak of the precision it computes in, times the reference's own
Fortran (compiled from its sources, oracle/_ref/ref_harness.exe) on a bounded
sample of the same workload on one host core, and measures the batched forms:
64 charge sets on one surface as a block, 48 molecules solved together, and
the autotuner's pick of leaf size, Taylor order and task granularity.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import json
import os
import sys
import time
from math import comb

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tabi-pb_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from tabipb import driver, hip, meshio, warmstart  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
METRIC = "PB solves/hour at ~500k faces, ΔG err<1e-3 vs FP64; GMRES iters learned vs 0"


def traffic_from_profiles(kernel, faces):
    """HBM bytes per launch of a kernel on a surface of `faces` faces, from the
    committed rocprofv3 --pmc summaries (profiles/*pmc*.md): their lines
    "traffic-bytes <kernel> <faces> <read B> <written B>" hold FETCH_SIZE and
    WRITE_SIZE per dispatch, scaled by the calibration the summary states.
    None when no summary covers this kernel and surface."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*pmc*.md"))):
        for ln in open(path):
            t = ln.strip().strip("`").split()
            if len(t) == 5 and t[0] == "traffic-bytes" and t[1] == kernel and int(t[2]) == faces:
                return float(t[3]) + float(t[4])
    return None


def w_far(p):
    """Algorithmic flops per accepted (target, cluster) pair."""
    return 20 * comb(p + 5, 3) + 64 * comb(p + 3, 3) + 40


def w_mom(p):
    return 9 * comb(p + 3, 3)


W_NEAR = 62


def workload(name):
    """Returns (driver.Problem, description dict)."""
    usr = meshio.UsrData(fname="mol", den="10", maxparnode=500, order=3, mac=0.8)
    if name.startswith("sphere"):
        sub = {"sphere20k": 5, "sphere82k": 6, "sphere328k": 7}[name]
        v, nv, f = meshio.icosphere(sub, 2.0)
        xyz, nrm, area = driver.face_geometry(v, nv, f)
        prob = driver.Problem(usr, xyz, nrm, area, np.zeros((1, 3)), np.ones(1), v, nv, f)
        desc = dict(workload=f"Born sphere a=2A q=+1, icosphere {len(f)} faces", faces=len(f),
                    atoms=1)
        return prob, desc
    if name.startswith("mol"):
        from tabipb import surface
        natom = int(name[3:].split("_")[0])
        h = 0.6 if natom >= 4000 else 0.5
        xyz, q, r = surface.synth_molecule(natom, seed=1)
        v, nv, f = surface.ses_mesh(xyz, r, probe=1.4, h=h)
        f, _ = driver.cull_faces(v, f)
        fxyz, nrm, area = driver.face_geometry(v, nv, f)
        # charges and positions as the driver reads them from a PQR (real*4)
        qpos = xyz.astype(np.float32).astype(np.float64)
        qq = q.astype(np.float32).astype(np.float64)
        prob = driver.Problem(usr, fxyz, nrm, area, qpos, qq, v, nv, f)
        desc = dict(workload=f"synthetic {natom}-atom molecule (seed 1), SES probe 1.4 A "
                             f"grid {h} A, {len(f)} faces", faces=len(f), atoms=natom)
        return prob, desc
    raise SystemExit(f"unknown workload {name}")


REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness.exe")


def stage_reference(prob, d):
    """The workload as a run directory of the reference program (usrdata.in,
    test_proteins/<fname>.{pqr,vert,face}), from the problem's own mesh."""
    usr = meshio.UsrData(**{**prob.usr.__dict__, "fname": "mol"})
    pref = os.path.join(d, "surf")
    meshio.write_msms(pref, prob.vpos, prob.vnrm, prob.faces)
    meshio.write_pqr(os.path.join(d, "mol.pqr"), prob.qpos, prob.q, np.full(len(prob.q), 1.5))
    meshio.stage_reference_run(d, usr, os.path.join(d, "mol.pqr"), pref)


def reference_sample(prob, kstride=64, lstride=16):
    """Runs the reference's own compiled routines (oracle/_ref/ref_harness.exe,
    built from the reference's Fortran sources) on a bounded sample of the
    workload, on one host core; returns the measured pieces."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        stage_reference(prob, d)
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec taskset -c 0 {REF_HARNESS} "
                                          f"sample={kstride} lstride={lstride}"],
                           cwd=d, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
    out = {}
    for ln in r.stdout.splitlines():
        t = ln.split()
        if len(t) == 5 and t[0] == "harness-sample":
            out[t[1]] = (float(t[2]), float(t[3]), float(t[4]))
        elif len(t) == 4 and t[0] == "harness-sample":
            out[t[1]] = (float(t[2]), float(t[3]))
    return out


def cpu_baseline(prob, iters):
    """Serial time of one reference solve of the workload, composed from the
    reference's own compiled routines timed on a bounded sample (one core):
    geometry, right-hand side and tree as measured (the right-hand side scaled
    from m of the charges), a matvec as the per-target traversal of every 64th
    target scaled to all targets plus the lazy moments of the first pass, a
    preconditioner application (LEAFMATVECpara refactors every leaf on every
    call) as the block fill and dgetrf/dgetrs of every 16th leaf scaled by the
    leaves' (2n)^2 and (2n)^3, and the energy from m charges; the solve is
    iters matvecs and iters+2 preconditioner applications (src/dgmres_dep.f),
    with iters the FP64 GPU solve's count (the same as the reference's on every
    fixture)."""
    t0 = time.perf_counter()
    m = reference_sample(prob)
    wall = time.perf_counter() - t0
    per = lambda k: m[k][0] * m[k][2] / m[k][1]      # noqa: E731
    t_first, t_second = m["matvec_first"][0], m["matvec"][0]
    t_mv = m["pb_kernel"][0] + (t_first - t_second) + per("matvec")
    t_ms = per("leaf_fill") + per("leaf_lu")
    t_setup = m["form_geom"][0] + per("form_rhs") + m["tree"][0]
    t_energy = per("energy")
    t_solve = t_setup + iters * t_mv + (iters + 2) * t_ms + t_energy
    lv = m["leaves"]
    sample = (f"reference Fortran (flang -O2, oracle/_ref/ref_harness.exe sample mode) on one "
              f"pinned core, {wall:.0f} s: every 64th matvec target "
              f"({int(m['matvec'][1])} of {int(m['matvec'][2])}), every 16th leaf block factored "
              f"and solved ({int(lv[0])} of {int(lv[1])}), right-hand side and energy from "
              f"{int(m['form_rhs'][1])} / {int(m['energy'][1])} of {int(m['energy'][2])} charges, "
              f"tree and geometry in full; solve = setup + {iters} matvecs + {iters + 2} "
              f"preconditioner applications + energy")
    return dict(value=3600.0 / t_solve, unit="solves/hour", cores=1, kind="reference",
                sample=sample, t_matvec_s=t_mv, t_msolve_s=t_ms, t_setup_s=t_setup,
                t_energy_s=t_energy, t_solve_s=t_solve)


def charge_set_surface(natom=1500, nsets=64):
    """The 1500-atom synthetic molecule (seed 1) and nsets charge vectors at its
    atom positions (N(0, 0.35) per atom, shifted to net -3 e); CPU only."""
    from tabipb import surface
    xyz, q, r = surface.synth_molecule(natom, seed=1)
    v, nv, f = surface.ses_mesh(xyz, r, probe=1.4, h=0.5)
    f, _ = driver.cull_faces(v, f)
    fx, fn, fa = driver.face_geometry(v, nv, f)
    qpos = xyz.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(7)
    Q = rng.normal(0.0, 0.35, (nsets, natom))
    Q -= (Q.sum(1, keepdims=True) + 3.0) / natom
    Q = Q.astype(np.float32).astype(np.float64)
    usr = meshio.UsrData(maxparnode=500, order=3, mac=0.8)
    return driver.Problem(usr, fx, fn, fa, qpos, Q[0]), Q


def batched_charge_sets(prob, Q, precision=1):
    """Many charge sets on one surface, solved as one block (one tree pass per
    GMRES step for all sets) and as single solves on the same set-up surface."""
    usr, nsets = prob.usr, Q.shape[0]
    S = hip.Surface(prob.params(precision), prob.xyz, prob.nrm, prob.area)
    qpos = prob.qpos
    S.solve_block(qpos, Q[:8], usr.epsp, want_x=False)       # warm-up
    t0 = time.perf_counter()
    blk = S.solve_block(qpos, Q, usr.epsp, want_x=False)
    t_blk = time.perf_counter() - t0
    t0 = time.perf_counter()
    its, sols = [], []
    for k in range(nsets):
        b = S.rhs(qpos, Q[k], usr.epsp)
        x, it, err, ierr = S.gmres(b, tol=1e-4)
        sols.append(S.energy(qpos, Q[k], usr.epsp, x)[0])
        its.append(it)
    t_seq = time.perf_counter() - t0
    S.close()
    sols = np.array(sols)
    n = prob.xyz.shape[0]
    return dict(workload=f"synthetic {prob.qpos.shape[0]}-atom molecule (seed 1), {n} faces, "
                         f"{nsets} charge sets", faces=int(n), sets=nsets,
                precision="mixed" if precision == 1 else "fp64",
                block_ms=t_blk * 1e3, sequential_ms=t_seq * 1e3, speedup=t_seq / t_blk,
                sets_per_s_block=nsets / t_blk, sets_per_s_sequential=nsets / t_seq,
                iters_equal=bool(np.array_equal(blk["iter"], its)),
                max_dG_rel_diff=float(np.max(np.abs(blk["soleng"] - sols) / np.abs(sols))),
                mean_iters=float(np.mean(its)))


def many_molecules(nmol=48, lo=200, hi=4000, workers=12):
    """nmol seeded synthetic molecules of lo..hi atoms (log-uniform), meshed in
    worker processes; CPU only."""
    seeds = list(range(700000, 700000 + nmol))
    sizes = warmstart.molecule_sizes(seeds, lo=lo, hi=hi, salt=2)
    return list(warmstart.molecules(seeds, sizes, workers=workers))


def packed_surfaces(mols, precision=1):
    """The molecules solved together (tabi_solve_many: one packed context for
    all surfaces -- per-surface trees, one list build, one batched leaf LU --
    and GMRES with its control on the device, one captured graph per restart
    cycle of every surface) and one after another."""
    prm = warmstart.params_for(precision)
    surfs = [(m["xyz"], m["nrm"], m["area"], m["qpos"], m["q"]) for m in mols]
    hip.solve_many(prm, surfs[:4], warmstart.EPS_P)             # warm-up
    t0 = time.perf_counter()
    res = hip.solve_many(prm, surfs, warmstart.EPS_P, tol=1e-4, graph=True)
    t_many = time.perf_counter() - t0
    t0 = time.perf_counter()
    single = [hip.solve(prm, *s, warmstart.EPS_P, tol=1e-4) for s in surfs]
    t_seq = time.perf_counter() - t0
    sol1 = np.array([r["soleng"] for r in single])
    faces = [int(m["xyz"].shape[0]) for m in mols]
    return dict(workload=f"{len(mols)} synthetic molecules of {min(m['natom'] for m in mols)}-"
                         f"{max(m['natom'] for m in mols)} atoms, {min(faces)}-{max(faces)} faces "
                         f"({sum(faces)} in total)",
                precision="mixed" if precision == 1 else "fp64",
                together_ms=t_many * 1e3, sequential_ms=t_seq * 1e3, speedup=t_seq / t_many,
                solves_per_hour_together=3600.0 * len(mols) / t_many,
                iters_equal=bool(np.array_equal(res["iter"], [r["iter"] for r in single])),
                max_dG_rel_diff=float(np.max(np.abs(res["soleng"] - sol1) / np.abs(sol1))),
                graph=res["stats"])


def autotuned(prob):
    """The autotuner's sweep (leaf 100-1000, p 2-6, two task granularities) on
    the charge-set surface, mixed precision, dG within 1e-3 of FP64."""
    from tabipb import autotune
    r = autotune.autotune(prob, task_exponents=(0, 4), repeats=1)
    default = [t for t in r["table"] if t["leaf"] == 500 and t["order"] == 3 and t["task_exp"] == 0][0]
    return dict(best=r["best"], default=default, speedup_vs_default=default["ms"] / r["best"]["ms"],
                candidates=len(r["table"]), rejected=sum(not t["ok"] for t in r["table"]),
                err_target=r["err_target"], reference=r["reference_form"],
                own_params_fp64_rel_err=r["own_params_rel_err"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="mol8000")
    ap.add_argument("--matvecs", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--kernels-only", type=int, default=0, metavar="K",
                    help="only K matvecs + preconditioner solves on the workload "
                         "(mixed precision), for counter collection")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the solve would shard as one all-gather of x per matvec).", file=sys.stderr)
        sys.exit(2)
    prob, desc = workload(a.workload)
    if a.kernels_only:
        S = hip.Surface(prob.params(precision=1), prob.xyz, prob.nrm, prob.area)
        dx = hip.DeviceArray.from_host(np.random.default_rng(0).standard_normal(2 * S.n))
        dy = hip.DeviceArray(2 * S.n)
        for _ in range(a.kernels_only):
            # in place: the library first copies x (7,737,152 B on the default
            # surface), the copy of known size the traffic counters are scaled by
            S.matvec_dev(dx.ptr, dx.ptr)
            S.msolve_dev(dx.ptr, dy.ptr)
        S.sync()
        print(f"[bench] {a.kernels_only} matvec + msolve on {S.n} faces", file=sys.stderr, flush=True)
        return
    # CPU preparation of the secondary workloads, before the GPU is used
    cs_prob = cs_Q = mols = None
    if not a.no_batch:
        cs_prob, cs_Q = charge_set_surface()
        mols = many_molecules()
    p = prob.params().order
    print(f"[bench] workload {desc['workload']} n={prob.xyz.shape[0]}", file=sys.stderr, flush=True)

    # inputs resident in device memory before the timed region (tabi_solve_dev)
    dprob = hip.DeviceProblem(prob.xyz, prob.nrm, prob.area, prob.qpos, prob.q)

    def time_solves(precision, model=None, steps=None, **over):
        # tabi_solve_dev returns after its last device work has completed
        prm = prob.params(precision, **over)
        steps = a.steps if steps is None else steps
        for _ in range(a.warmup):
            r = dprob.solve(prm, prob.usr.epsp, tol=1e-4, model=model)
        t0 = time.perf_counter()
        for _ in range(steps):
            r = dprob.solve(prm, prob.usr.epsp, tol=1e-4, model=model)
        dt = (time.perf_counter() - t0) / steps
        print(f"[bench] precision={precision} learned={model is not None} {over} solve "
              f"{dt * 1e3:.1f} ms iter={r['iter']} ierr={r['ierr']} dG={r['soleng']:.6f} "
              f"phases={ {k: round(float(v), 2) for k, v in r['times_ms'].items()} }",
              file=sys.stderr, flush=True)
        return dt, r

    dt64, res64 = time_solves(0)
    dt32, res32 = time_solves(1)
    model = hip.WsModel(warmstart.load()) if os.path.exists(warmstart.DEFAULT_MODEL) else None
    dtws, resws = time_solves(1, model) if model is not None else (None, None)
    # the configuration's other expansion: p = 6, theta = 0.5
    p6 = {prec_: time_solves(prec_, steps=max(1, a.steps // 2), order=6, theta=0.5)
          for prec_ in (0, 1)}
    forms = {"fp64": (dt64, res64, 0), "mixed": (dt32, res32, 1)}
    if model is not None:
        forms["mixed+learned"] = (dtws, resws, 1)
    ok = {k: v for k, v in forms.items()
          if v[1]["ierr"] == 0 and abs(v[1]["soleng"] - res64["soleng"]) <= 1e-3 * abs(res64["soleng"])}
    form = min(ok, key=lambda k: ok[k][0])
    dt, res, prec = forms[form]
    derr = abs(res["soleng"] - res64["soleng"]) / abs(res64["soleng"])
    # the same form through the host-pointer entry point (geometry and charges
    # uploaded, solution downloaded inside each solve): PCIe-inclusive time
    fmodel = model if form == "mixed+learned" else None
    driver.solve(prob, tol=1e-4, precision=prec, model=fmodel)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        driver.solve(prob, tol=1e-4, precision=prec, model=fmodel)
    dt_host = (time.perf_counter() - t0) / a.steps
    dprob.free()

    def kernel_profile(precision):
        S = hip.Surface(prob.params(precision=precision), prob.xyz, prob.nrm, prob.area)
        st = S.stats()
        n = S.n
        dx = hip.DeviceArray.from_host(np.random.default_rng(0).standard_normal(2 * n))
        dy = hip.DeviceArray(2 * n)
        S.matvec_dev(dx.ptr, dy.ptr)
        S.sync()
        S.profile(True)
        for _ in range(a.matvecs):
            S.matvec_dev(dx.ptr, dy.ptr)
        kt = S.kernel_times()
        S.profile(False)
        S.sync()
        t1 = time.perf_counter()
        for _ in range(a.matvecs):
            S.matvec_dev(dx.ptr, dy.ptr)
        S.sync()
        mv_ms = (time.perf_counter() - t1) / a.matvecs * 1e3
        S.profile(True)
        for _ in range(a.matvecs):
            S.msolve_dev(dy.ptr, dx.ptr)
        ms_t = S.kernel_times()["msolve"]
        S.profile(False)
        S.close()
        dx.free()
        dy.free()
        return st, kt, mv_ms, ms_t[0] / max(ms_t[1], 1)

    st, kt, mv_ms, msolve_ms = kernel_profile(prec)
    _, kt64, mv64_ms, _ = kernel_profile(0) if prec == 1 else (st, kt, mv_ms, msolve_ms)
    flops = {"near": W_NEAR * st["p_near"], "far": w_far(p) * st["p_far"],
             "moments": w_mom(p) * st["p_mom"]}
    avg = {k: kt[k][0] / max(kt[k][1], 1) for k in ("near", "far", "moments", "reduce")}
    dom = max(("near", "far"), key=lambda k: avg[k])
    achieved = flops[dom] / (avg[dom] * 1e-3) / 1e12
    kname = "k_" + dom + ("32" if prec == 1 else "") + ("<%d>" % p if dom == "far" else "")
    peak = FP32_PEAK_TFLOPS if prec == 1 else FP64_PEAK_TFLOPS
    print(f"[bench] matvec {mv_ms:.3f} ms (fp64 {mv64_ms:.3f} ms); kernel avg ms {avg}; "
          f"msolve {msolve_ms:.3f} ms; stats {st}", file=sys.stderr, flush=True)

    batch = packed = tuned = None
    if not a.no_batch:
        batch = batched_charge_sets(cs_prob, cs_Q)
        print(f"[bench] batched charge sets {batch}", file=sys.stderr, flush=True)
        packed = packed_surfaces(mols)
        print(f"[bench] many surfaces {packed}", file=sys.stderr, flush=True)
        tuned = autotuned(cs_prob)
        print(f"[bench] autotune {tuned}", file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(prob, res64["iter"])
        print(f"[bench] cpu baseline {cpu}", file=sys.stderr, flush=True)

    value = 3600.0 / dt
    out = {
        "metric": METRIC, "value": value, "unit": "solves/hour", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f32+f64+bf16" if form == "mixed+learned" else "f32+f64") if prec == 1
                 else "f64", "data": "synthetic",
        "config": dict(desc, leaf=prob.usr.maxparnode, order=p, theta=prob.usr.mac, tol=1e-4,
                       form={"fp64": "FP64, zero start",
                             "mixed": "mixed precision (FP32 near/far field, compensated "
                                      "sums; FP64 GMRES, leaf LU, self term, energy), zero start",
                             "mixed+learned": "mixed precision, learned initial guess "
                                              "(BF16 MLP)"}[form],
                       gmres_iter=res["iter"], gmres_iter_zero_start=res32["iter"],
                       gmres_iter_learned_start=None if resws is None else resws["iter"],
                       learned_guess_used=None if resws is None else resws["ws_used"],
                       gmres_iter_fp64_zero_start=res64["iter"],
                       soleng_kcal=res["soleng"], soleng_fp64_kcal=res64["soleng"],
                       dG_rel_err_vs_fp64=derr,
                       fp64_solves_per_hour=3600.0 / dt64, fp64_ms_per_solve=dt64 * 1e3,
                       mixed_solves_per_hour=3600.0 / dt32, mixed_ms_per_solve=dt32 * 1e3,
                       mixed_learned_solves_per_hour=None if dtws is None else 3600.0 / dtws,
                       mixed_learned_ms_per_solve=None if dtws is None else dtws * 1e3,
                       pcie_inclusive_ms_per_solve=dt_host * 1e3,
                       p6_theta05={k: dict(ms_per_solve=v[0] * 1e3, solves_per_hour=3600.0 / v[0],
                                           gmres_iter=v[1]["iter"], soleng_kcal=v[1]["soleng"],
                                           dG_rel_diff_vs_p3_fp64=abs(v[1]["soleng"] - res64["soleng"])
                                           / abs(res64["soleng"]))
                                   for k, v in (("fp64", p6[0]), ("mixed", p6[1]))},
                       matvec_ms=mv_ms, fp64_matvec_ms=mv64_ms, msolve_ms=msolve_ms,
                       matvecs_per_s=1e3 / mv_ms,
                       phases_ms={k: float(v) for k, v in res["times_ms"].items()},
                       batched_charge_sets=batch, many_surfaces=packed, autotune=tuned),
        # compute-bound on the vector ALUs (exp/rsqrt pair sums, no matrix
        # cores): priced against the published FP32 (or FP64) vector peak; the
        # kernels are built without packed FP32, whose issue rate is half the
        # FP32 figure
        "roofline": {"bound": "valu", "kernel": kname,
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "traffic": traffic_from_profiles(kname, int(prob.xyz.shape[0])),
                     "alg_flops_per_launch": flops[dom], "avg_launch_ms": avg[dom]},
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in
                                                  ("value", "unit", "cores", "kind", "sample")},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
