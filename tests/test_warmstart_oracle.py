"""The warm-start oracle pinned against mathematics: BF16 rounding against its
definition, the GELU derivative and the FP64 backward pass against finite
differences, the curvature feature against the exact 1/R of a sphere and the
nearest-charge distances against a k-d tree.  Also checks that the product
module's packed layout agrees with the oracle's."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import ws_oracle as W
from tabipb import driver, meshio
from tabipb import warmstart as WS


def test_bf16_rounding_known_values_and_bound():
    one = np.float32(1.0)
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20,
                     -2.5, 3.0e-30, 65504.0], dtype=np.float32)
    got = W.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1) wins;
    # 1 + 3*2^-8 ties between 1 + 2^-7 (odd) and 1 + 2^-6 (even): 1 + 2^-6
    assert got[0] == one
    assert got[1] == one
    assert got[2] == np.float32(1.0 + 2.0 ** -6)
    assert got[3] == np.float32(1.0 + 2.0 ** -7)
    assert got[4] == np.float32(-2.5)
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(100000) * 10.0 ** rng.uniform(-20, 20, 100000)).astype(np.float32)
    r = W.bf16(x)
    # representable: low 16 bits zero; error at most half a BF16 ulp (2^-8 relative)
    assert np.all(r.view(np.uint32) & 0xFFFF == 0)
    assert np.all(np.abs(r.astype(np.float64) - x) <= np.abs(x) * 2.0 ** -8 * (1 + 1e-6))
    # idempotent
    assert np.array_equal(W.bf16(r), r)


def test_gelu_derivative_matches_finite_differences():
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (W.gelu(z + h) - W.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - W.gelu_d(z))) < 1e-8
    # GELU limits
    assert abs(W.gelu(10.0) - 10.0) < 1e-12 and abs(W.gelu(-10.0)) < 1e-12


def _random_model(seed, nsamp=40):
    rng = np.random.default_rng(seed)
    feat = rng.standard_normal((nsamp, 12)) * 2 + 1
    tgt = rng.standard_normal((nsamp, 2))
    P = WS.init_params(feat.mean(0), feat.std(0), np.array([0.7, 1.3]), seed).astype(np.float64)
    P[W.O_B1:W.O_W2] = rng.standard_normal(128) * 0.1
    P[W.O_B4:W.NP] = rng.standard_normal(2) * 0.1
    P[W.O_W4:W.O_B4] *= 5.0
    return P, feat, tgt


def test_fp64_backward_matches_finite_differences():
    P, feat, tgt = _random_model(3)
    g, loss = W.loss_grad(P, feat, tgt, emulate=False)
    assert loss == pytest.approx(W.loss_only(P, feat, tgt), rel=1e-14)
    rng = np.random.default_rng(4)
    blocks = [(W.O_W1, W.O_B1), (W.O_B1, W.O_W2), (W.O_W2, W.O_B2), (W.O_B2, W.O_W3),
              (W.O_W3, W.O_B3), (W.O_B3, W.O_W4), (W.O_W4, W.O_B4), (W.O_B4, W.NP)]
    h = 1e-6
    for a, b in blocks:
        for k in rng.choice(np.arange(a, b), size=min(6, b - a), replace=False):
            Pp, Pm = P.copy(), P.copy()
            Pp[k] += h
            Pm[k] -= h
            fd = (W.loss_only(Pp, feat, tgt) - W.loss_only(Pm, feat, tgt)) / (2 * h)
            assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(fd)) + 1e-9, (k, fd, g[k])
    # the normalisation header is not trained
    assert np.all(g[:W.O_W1] == 0)


def test_emulated_gradient_close_to_fp64():
    P, feat, tgt = _random_model(5, nsamp=256)
    g64, l64 = W.loss_grad(P, feat, tgt, emulate=False)
    g16, l16 = W.loss_grad(P, feat, tgt, emulate=True)
    assert l16 == pytest.approx(l64, rel=2e-2)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 3e-2


def test_packed_layout_agrees():
    assert WS.n_params() == W.NP
    P = WS.init_params(np.zeros(12), np.ones(12), np.ones(2), 0)
    u = W.unpack(P)
    assert u["W1"].shape == (128, 12) and u["W4"].shape == (2, 128)
    assert np.all(u["b1"] == 0) and np.all(u["sig"] == 1)


def test_curvature_feature_is_inverse_radius_on_sphere():
    R = 3.0
    v, nv, f = meshio.icosphere(5, R)
    xyz, nrm, area = driver.face_geometry(v, nv, f)
    n = xyz.shape[0]
    # leaves: the whole sphere as 20 contiguous blocks (any partition works)
    order = np.lexsort((xyz[:, 2], xyz[:, 1], xyz[:, 0]))
    xyz, nrm, area = xyz[order], nrm[order], area[order]
    edges = np.linspace(0, n, 21).astype(int)
    leaves = [(edges[i], edges[i + 1] - 1) for i in range(20)]
    b = np.ones(2 * n)
    feat, s = W.features(xyz, nrm, area, leaves, np.zeros((1, 3)), np.ones(1), 1.0, b)
    assert s == pytest.approx(1.0)
    k = feat[:, 7]
    # centroids sit slightly inside the sphere and normals are vertex-normal
    # averages: the estimate is 1/R to within a percent on average
    assert np.median(k) == pytest.approx(1.0 / R, rel=1e-2)
    r = np.linalg.norm(xyz, axis=1)
    assert np.all(np.abs(feat[:, 8] - r) < 1e-5)
    # one unit charge at the centre: windowed potentials are exactly
    # w(r)/(4 pi r), the normal field w(r)/(4 pi r^2) pointing inward (cos = -1)
    np.testing.assert_allclose(feat[:, 2], np.exp(-r ** 2 / 25.0) / (4 * np.pi * r), rtol=1e-5)
    np.testing.assert_allclose(feat[:, 5], -np.exp(-r ** 2 / 100.0) / (4 * np.pi * r ** 2),
                               rtol=2e-3)
    np.testing.assert_allclose(feat[:, 10], 1.0, atol=2e-3)
    np.testing.assert_allclose(feat[:, 11], np.exp(-r ** 2 / 100.0), rtol=1e-5)


def test_nearest_charge_distances_match_kdtree():
    rng = np.random.default_rng(9)
    c = rng.uniform(-10, 10, (500, 3))
    q = rng.uniform(-10, 10, (300, 3))
    feat, _ = W.features(c, np.tile([0.0, 0.0, 1.0], (500, 1)), np.ones(500), [(0, 499)], q,
                         np.ones(300), 1.0, np.ones(1000))
    d, idx = cKDTree(q).query(c, k=2)
    np.testing.assert_allclose(feat[:, 8:10], d, rtol=1e-5)
    cosv = (c[:, 2] - q[idx[:, 0], 2]) / d[:, 0]
    np.testing.assert_allclose(feat[:, 10], cosv, atol=1e-5)
