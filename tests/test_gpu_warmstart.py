"""GPU warm start: the BF16 MFMA inference and training kernels against the
CPU oracle (same BF16 rounding points emulated), the feature kernel against the
oracle's features, training progress, and the end-to-end solve: GMRES from the
learned guess reaches the same tolerance in fewer iterations, and falls back to
the zero start when the guess is worse than none.

Tolerances: predictions within 1e-2 relative L2 of the emulating oracle (BF16
rounding of near-equal FP32 values can flip a last bit), gradients within 2e-2
relative L2 per parameter block, the loss within 1e-3; features within 1e-5
relative (FP32 sums), the windowed sums within 1e-4 of their largest value;
the solvation energy of the warm-started solve
within 1e-3 of the FP64 solve.
"""
import numpy as np
import pytest

import oracle as O
import ws_oracle as W
from tabipb import driver, hip, surface
from tabipb import warmstart as WS

pytestmark = pytest.mark.gpu

BLOCKS = {"W1": (W.O_W1, W.O_B1), "b1": (W.O_B1, W.O_W2), "W2": (W.O_W2, W.O_B2),
          "b2": (W.O_B2, W.O_W3), "W3": (W.O_W3, W.O_B3), "b3": (W.O_B3, W.O_W4),
          "W4": (W.O_W4, W.O_B4), "b4": (W.O_B4, W.NP)}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300))


def _model(seed, ns):
    rng = np.random.default_rng(seed)
    feat = (rng.standard_normal((ns, 12)) * [1, 2, 1, 1, 1, .1, .3, 2, 2, 2, 2, 2]
            + [0, 0, 0, 0, 0, .1, .2, 3, 4, 5, 6, 7]).astype(np.float32)
    tgt = rng.standard_normal((ns, 2)).astype(np.float32)
    P = WS.init_params(feat.mean(0), feat.std(0) + 0.25, np.array([0.8, 1.7]), seed)
    P[W.O_B1:W.O_W2] = rng.standard_normal(128) * 0.1
    P[W.O_B2:W.O_W3] = rng.standard_normal(128) * 0.1
    P[W.O_W4:W.O_B4] *= 5.0
    return P.astype(np.float32), feat, tgt


@pytest.mark.parametrize("ns", [1, 31, 32, 1000, 40000])
def test_predict_matches_oracle(ns):
    print(f"[gpu] ws predict ns={ns}", flush=True)
    P, feat, _ = _model(1, ns)
    M = hip.WsModel(P)
    x = M.predict(feat, 2.5)
    xo = W.predict(P, feat, 2.5, emulate=True)
    assert rel(x, xo) < 1e-2
    x64 = W.predict(P, feat, 2.5, emulate=False)
    assert rel(x, x64) < 5e-2
    M.close()


@pytest.mark.parametrize("ns", [7, 32, 4096, 20000])
def test_gradient_matches_oracle(ns):
    print(f"[gpu] ws grad ns={ns}", flush=True)
    P, feat, tgt = _model(2, ns)
    M = hip.WsModel(P)
    g, loss = M.grad(feat, tgt)
    go, lo = W.loss_grad(P, feat, tgt, emulate=True)
    assert loss == pytest.approx(lo, rel=1e-3)
    for name, (a, b) in BLOCKS.items():
        assert rel(g[a:b], go[a:b]) < 2e-2, name
    assert np.all(g[:W.O_W1] == 0)
    M.close()


def test_training_reduces_loss_on_a_learnable_target():
    print("[gpu] ws train synthetic", flush=True)
    rng = np.random.default_rng(3)
    ns = 50000
    feat = rng.standard_normal((ns, 12)).astype(np.float32)
    tgt = np.stack([np.tanh(feat[:, 0] + 0.5 * feat[:, 7]), feat[:, 1] * feat[:, 2]], 1)
    tgt = tgt.astype(np.float32)
    P = WS.init_params(np.zeros(12), np.ones(12), np.ones(2), 0)
    M = hip.WsModel(P)
    losses = M.train(feat, tgt, epochs=8, batch=4096, lr=3e-3, seed=1)
    assert losses[-1] < 0.2 * losses[0]
    # the trained model's own loss agrees with the last epoch's
    g, loss = M.grad(feat, tgt)
    assert loss < 1.5 * losses[-1]
    M.close()


def _mol(seed, natom, h=0.6):
    return WS.molecule(seed, natom, h=h)


def test_features_match_oracle():
    m = _mol(21, 120)
    print(f"[gpu] ws features faces={m['xyz'].shape[0]}", flush=True)
    S = hip.Surface(WS.params_for(0), m["xyz"], m["nrm"], m["area"])
    feat, s = S.ws_features(m["qpos"], m["q"], WS.EPS_P)
    b = S.rhs(m["qpos"], m["q"], WS.EPS_P)
    perm = S.perm
    txyz, order, nodes = O.build_tree(m["xyz"], WS.LEAF)
    assert np.array_equal(order, perm)
    ib, ie = O.leaves(nodes)
    fo, so = W.features(m["xyz"][perm], m["nrm"][perm], m["area"][perm], list(zip(ib, ie)),
                        m["qpos"], m["q"], WS.EPS_P, b)
    assert s == pytest.approx(so, rel=1e-12)
    for k in range(12):
        tol = 1e-4 if k in (2, 3, 4, 5, 11) else 1e-5
        np.testing.assert_allclose(feat[:, k], fo[:, k], rtol=tol,
                                   atol=tol * np.abs(fo[:, k]).max(), err_msg=str(k))
    S.close()


def test_learned_start_cuts_iterations_and_falls_back():
    """A small model trained on six molecules, then a held-out one."""
    train = [_mol(500 + k, 150 + 30 * k) for k in range(6)]
    rng = np.random.default_rng(0)
    F, T = [], []
    for m in train:
        f, t, n, it = WS.samples(m, 4096, rng)
        print(f"[gpu] ws data faces={n} iters={it}", flush=True)
        F.append(f)
        T.append(t)
    P, losses = WS.fit(np.concatenate(F), np.concatenate(T), epochs=30, batch=4096, lr=2e-3,
                       log=None)
    assert losses[-1] < 0.5 * losses[0]
    m = _mol(777, 260)
    a = (m["xyz"], m["nrm"], m["area"], m["qpos"], m["q"], WS.EPS_P)
    M = hip.WsModel(P)
    r64 = hip.solve(WS.params_for(0), *a, tol=1e-6)
    r0 = hip.solve(WS.params_for(0), *a, tol=1e-4)
    r1 = hip.solve(WS.params_for(0), *a, tol=1e-4, model=M)
    print(f"[gpu] ws held-out iters zero {r0['iter']} learned {r1['iter']} "
          f"used {r1['ws_used']}", flush=True)
    assert r1["ierr"] == 0 and r1["err"] <= 1e-4
    assert r1["ws_used"] == 1
    assert r1["iter"] < r0["iter"]
    assert abs(r1["soleng"] - r64["soleng"]) / abs(r64["soleng"]) < 1e-3
    # mixed precision with the learned start
    r2 = hip.solve(WS.params_for(1), *a, tol=1e-4, model=M)
    assert r2["ierr"] == 0 and r2["iter"] < r0["iter"]
    assert abs(r2["soleng"] - r64["soleng"]) / abs(r64["soleng"]) < 1e-3
    # a guess far worse than zero: the zero start is taken, same iterations
    Pbad = P.copy()
    Pbad[W.O_B4:W.NP] = 1.0e4
    Mb = hip.WsModel(Pbad)
    r3 = hip.solve(WS.params_for(0), *a, tol=1e-4, model=Mb)
    assert r3["ws_used"] == 0 and r3["iter"] == r0["iter"] and r3["ierr"] == 0
    assert r3["soleng"] == pytest.approx(r0["soleng"], rel=1e-10)
    # the surface-level entry point agrees with the full solve
    S = hip.Surface(WS.params_for(0), m["xyz"], m["nrm"], m["area"])
    b = S.rhs(m["qpos"], m["q"], WS.EPS_P)
    x, it, err, ierr, used = S.gmres_ws(M, m["qpos"], m["q"], WS.EPS_P, b, tol=1e-4)
    assert it == r1["iter"] and used == 1 and ierr == 0
    S.close()
    M.close()
    Mb.close()


def test_model_api_errors():
    P = WS.init_params(np.zeros(12), np.ones(12), np.ones(2), 0)
    with pytest.raises(hip.TabiError, match="parameters"):
        hip.WsModel(P[:-1])
    bad = P.copy()
    bad[12] = 0.0
    with pytest.raises(hip.TabiError, match="positive"):
        hip.WsModel(bad)
    M = hip.WsModel(P)
    with pytest.raises(hip.TabiError, match="shape"):
        M.predict(np.zeros((4, 11)), 1.0)
    np.testing.assert_array_equal(M.params(), P)
    M.close()


def test_committed_model_gains_on_held_out_molecules():
    """The shipped model (trained on 1024 seeded molecules of 300-3000 atoms)
    on held-out seeds: GMRES from its guess converges to the same tolerance in
    fewer iterations, with the energy within 1e-3 of the FP64 solve."""
    P = WS.load()
    seeds = list(range(900000, 900004))
    sizes = WS.molecule_sizes(seeds, salt=1)
    rows = WS.evaluate(P, seeds, sizes, workers=1, log=lambda s: print(s, flush=True))
    it0, it1, used = rows[:, 1], rows[:, 2], rows[:, 3]
    err = np.abs(rows[:, 5] - rows[:, 6]) / np.abs(rows[:, 6])
    assert np.all(used == 1)
    assert it1.sum() < it0.sum()
    assert np.all(it1 <= it0)
    assert np.all(err < 1e-3)
