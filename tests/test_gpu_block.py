"""Blocks of right-hand sides on one surface (one tree pass for many vectors)
against the single-vector entry points, column by column.

Per column the block kernels repeat the single kernels' arithmetic, so the
tolerances are tight: matvec and preconditioner columns within 1e-14 relative
L2 of the single results (the compiler may contract a product into a fused
multiply-add differently in the two kernels), GMRES runs take the same number
of iterations and agree to 1e-10, and the energies of a block solve agree with
single solves to 1e-10.
"""
import numpy as np
import pytest

import cases
from tabipb import driver, hip
from tabipb import warmstart as WS

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300))


def _surface(name, precision):
    v, nv, f, atoms, usr = cases.build(name)
    f2, _ = driver.cull_faces(v, f)
    xyz, nrm, area = driver.face_geometry(v, nv, f2)
    P = driver.Problem(usr, xyz, nrm, area, atoms[:, :3], atoms[:, 3])
    return P, hip.Surface(P.params(precision=precision), xyz, nrm, area)


@pytest.mark.parametrize("name,precision", [("sph3_l100_p3", 0), ("sph3_l100_p3", 1),
                                            ("sph4_off_p6", 0), ("sph4_off_p6", 1),
                                            ("sph3_k0_p2", 0)])
def test_block_matvec_and_msolve_match_single(name, precision):
    print(f"[gpu] block matvec {name} precision={precision}", flush=True)
    P, S = _surface(name, precision)
    rng = np.random.default_rng(11)
    X = rng.standard_normal((11, 2 * S.n))
    X[4] = 0.0                     # the zero-vector shortcut applies per column
    Y = S.matvec_block(X)
    Z = S.msolve_block(X)
    for k in range(X.shape[0]):
        y1 = S.matvec(X[k])
        z1 = S.msolve(X[k])
        if k == 4:
            assert np.array_equal(Y[k], X[k]) and np.array_equal(y1, X[k])
            assert np.all(Z[k] == 0)
            continue
        assert rel(Y[k], y1) <= 1e-14, k
        assert rel(Z[k], z1) <= 1e-14, k
    # a block of one column
    assert rel(S.matvec_block(X[:1])[0], S.matvec(X[0])) <= 1e-14
    S.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_block_gmres_matches_single_runs(precision):
    print(f"[gpu] block gmres precision={precision}", flush=True)
    P, S = _surface("sph3_l100_p3", precision)
    rng = np.random.default_rng(5)
    B = rng.standard_normal((6, 2 * S.n))
    B[2] *= 1e-3                    # scale does not matter to the stopping rule
    X0 = np.zeros_like(B)
    X0[3] = 0.1 * rng.standard_normal(2 * S.n)   # a nonzero initial guess
    X, it, err, ierr = S.gmres_block(B, X0, tol=1e-6)
    for k in range(B.shape[0]):
        x1, it1, err1, ierr1 = S.gmres(B[k], X0[k], tol=1e-6)
        assert it[k] == it1 and ierr[k] == ierr1 == 0, (k, it[k], it1)
        assert rel(X[k], x1) <= 1e-10
        assert err[k] == pytest.approx(err1, rel=1e-8)
    S.close()


def test_block_solve_of_charge_sets_matches_single_solves():
    """A synthetic molecule with 10 charge sets at fixed atom positions
    (charges resampled per set, net charge -3), as one block against ten
    single solves."""
    m = WS.molecule(31, 200, h=0.6)
    print(f"[gpu] block solve faces={m['xyz'].shape[0]}", flush=True)
    rng = np.random.default_rng(2)
    nq = m["q"].shape[0]
    Q = rng.normal(0.0, 0.35, (10, nq))
    Q -= (Q.sum(1, keepdims=True) + 3.0) / nq
    Q = Q.astype(np.float32).astype(np.float64)
    for precision in (0, 1):
        S = hip.Surface(WS.params_for(precision), m["xyz"], m["nrm"], m["area"])
        res = S.solve_block(m["qpos"], Q, WS.EPS_P, tol=1e-4)
        for k in range(Q.shape[0]):
            b = S.rhs(m["qpos"], Q[k], WS.EPS_P)
            x1, it1, err1, ierr1 = S.gmres(b, tol=1e-4)
            sol1, cou1 = S.energy(m["qpos"], Q[k], WS.EPS_P, x1)
            assert res["iter"][k] == it1 and res["ierr"][k] == 0
            assert rel(res["x"][k], x1) <= 1e-10
            assert res["soleng"][k] == pytest.approx(sol1, rel=1e-10)
            assert res["couleng"][k] == pytest.approx(cou1, rel=1e-12)
        S.close()


def test_block_api_errors():
    P, S = _surface("sph3_l100_p3", 0)
    with pytest.raises(hip.TabiError, match="shape"):
        S.matvec_block(np.zeros((2, 2 * S.n + 1)))
    with pytest.raises(hip.TabiError, match="shape"):
        S.solve_block(np.zeros((3, 3)), np.zeros((2, 4)), 1.0)
    S.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_many_surfaces_match_single_solves(precision):
    """Five molecules of uneven size solved together (lockstep GMRES, each
    step's matvec + preconditioner solve of all surfaces replayed from one
    captured graph) against five single solves: same iterations, energies to
    1e-10.  The graph is replayed once per GMRES step and captured again when a
    surface drops out, so every replay is checked through the final energies;
    the eager form (no graph) must agree too."""
    mols = [WS.molecule(40 + k, na, h=0.6) for k, na in enumerate((60, 150, 90, 220, 120))]
    surfs = [(m["xyz"], m["nrm"], m["area"], m["qpos"], m["q"]) for m in mols]
    print(f"[gpu] many surfaces faces={[m['xyz'].shape[0] for m in mols]}", flush=True)
    prm = WS.params_for(precision)
    res = hip.solve_many(prm, surfs, WS.EPS_P, tol=1e-4, graph=True, packed=False)
    eager = hip.solve_many(prm, surfs, WS.EPS_P, tol=1e-4, graph=False, packed=False)
    pk = hip.solve_many(prm, surfs, WS.EPS_P, tol=1e-4)
    assert res["stats"]["graph_launches"] > 0 and res["stats"]["captures"] >= 1
    for k, sfc in enumerate(surfs):
        r1 = hip.solve(prm, *sfc, WS.EPS_P, tol=1e-4)
        # packed: every surface's arithmetic is its own context's, bit for bit
        assert pk["iter"][k] == r1["iter"] and pk["ierr"][k] == 0
        assert pk["err"][k] == r1["err"]
        assert pk["soleng"][k] == r1["soleng"] and pk["couleng"][k] == r1["couleng"]
        assert res["iter"][k] == r1["iter"] and res["ierr"][k] == 0
        assert res["soleng"][k] == pytest.approx(r1["soleng"], rel=1e-10)
        assert res["couleng"][k] == pytest.approx(r1["couleng"], rel=1e-12)
        assert eager["iter"][k] == r1["iter"]
        assert eager["soleng"][k] == pytest.approx(r1["soleng"], rel=1e-10)


@pytest.mark.parametrize("precision", [0, 1])
def test_packed_context_matvec_and_msolve_equal_own_contexts(precision):
    """Four molecules packed into one context (one tree and list set per
    surface, one launch per kernel for all): every surface's entries of the
    packed product and preconditioner solve equal its own context's, bit for
    bit, and the padding between surfaces stays zero."""
    mols = [WS.molecule(60 + k, na, h=0.6) for k, na in enumerate((90, 40, 170, 60))]
    print(f"[gpu] packed faces={[m['xyz'].shape[0] for m in mols]} precision={precision}", flush=True)
    prm = WS.params_for(precision)
    P = hip.PackedSurfaces(prm, [(m["xyz"], m["nrm"], m["area"]) for m in mols])
    assert np.all(P.face0 % 64 == 0) and P.face0[0] == 0
    assert np.all(np.diff(P.face0) >= P.nface)
    rng = np.random.default_rng(3)
    xs = [rng.standard_normal(2 * m["xyz"].shape[0]) for m in mols]
    xp = P.pack(xs)
    yp, zp = P.matvec(xp), P.msolve(xp)
    Y, Z = P.unpack(yp), P.unpack(zp)
    pad = np.ones(2 * P.n, dtype=bool)
    for k, m in enumerate(mols):
        f0, nk = int(P.face0[k]), int(P.nface[k])
        pad[f0:f0 + nk] = False
        pad[P.n + f0:P.n + f0 + nk] = False
        S = hip.Surface(prm, m["xyz"], m["nrm"], m["area"])
        assert np.array_equal(Y[k], S.matvec(xs[k])), k
        assert np.array_equal(Z[k], S.msolve(xs[k])), k
        S.close()
    assert np.all(yp[pad] == 0.0)
    P.close()
