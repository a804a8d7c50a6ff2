"""Pins of the learned-FSAI oracle (CPU only).

Nothing in the reference covers the learned preconditioner, so each oracle
function is checked against something it does not restate: finite
differences, hand-computed values, closed forms.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import clsparse_amd as cls
from oracle import fsai as of

_M64 = (1 << 64) - 1


def test_bf16_round_to_nearest_even():
    one = 1.0
    cases = {
        one: one,
        one + 2.0 ** -8: one,                       # tie, 1.0 has an even last bit
        one + 2.0 ** -8 + 2.0 ** -20: one + 2.0 ** -7,
        one + 3 * 2.0 ** -8: one + 2.0 ** -6,       # tie, rounds up to the even neighbour
        -2.5: -2.5,
        0.0: 0.0,
        3.0e-3: 0.0030059814453125,                 # 0x3B45 -> 0.0030059814453125
    }
    got = of.bf16_round(np.array(list(cases), dtype=np.float32))
    np.testing.assert_array_equal(got, np.array(list(cases.values()), dtype=np.float32))


def _splitmix_finaliser(x):
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def test_rademacher_matches_scalar_integer_arithmetic():
    seed = of.probe_seed(12345, 3)
    assert seed == (12345 * 0x100000001B3 + 3) & _M64
    z = of.rademacher(50, seed)
    for i in range(50):
        x = _splitmix_finaliser(seed ^ ((0x9E3779B97F4A7C15 * (i + 1)) & _M64))
        assert z[i] == (1.0 if x & 1 else -1.0)
    big = of.rademacher(200000, 7)
    assert set(np.unique(big)) == {-1.0, 1.0}
    assert abs(big.mean()) < 0.01


def _tiny():
    a = np.array([[4.0, -1.0, 0.0], [-1.0, 4.0, -2.0], [0.0, -2.0, 9.0]])
    m = sp.csr_matrix(a)
    return m.indptr.astype(np.uint32), m.indices.astype(np.uint32), m.data, a


def test_features_hand_computed():
    rp, ci, v, _ = _tiny()
    X = of.features(rp, ci, v)
    # tril pattern: (0,0) (1,0) (1,1) (2,1) (2,2); dhat = 1/2, 1/2, 1/3
    a = [1.0, -0.25, 1.0, -1.0 / 3.0, 1.0]
    r = [0.25, 0.25 + 1.0 / 3.0, 1.0 / 3.0]
    rows = [0, 1, 1, 2, 2]
    cols = [0, 0, 1, 1, 2]
    lens = [2, 3, 2]
    for k in range(5):
        i, j = rows[k], cols[k]
        exp = [a[k], float(i == j), r[i], r[j], -r[i], -r[j], np.float32(a[k]) ** 2, lens[i] / 16]
        np.testing.assert_allclose(X[k, :8], np.array(exp, dtype=np.float32), rtol=1e-7)
        assert np.all(X[k, 8:] == 0)


def test_assemble_g_hand_computed():
    rp, ci, v, _ = _tiny()
    G = of.assemble_g(rp, ci, v).toarray()
    np.testing.assert_allclose(G, np.diag([0.5, 0.5, 1.0 / 3.0]), rtol=0, atol=0)
    o = np.array([0.1, -0.2, 0.3, 0.4, -0.5])
    G = of.assemble_g(rp, ci, v, o).toarray()
    exp = np.array([[1.1 * 0.5, 0, 0], [-0.2 * 0.5, 1.3 * 0.5, 0], [0, 0.4 / 3.0, 0.5 / 3.0]])
    np.testing.assert_allclose(G, exp, rtol=1e-15)


def test_hutchinson_zero_for_diagonal_matrix_with_jacobi_factor():
    d = np.array([2.0, 5.0, 0.25, 9.0])
    A = sp.diags(d).tocsr()
    G = of.assemble_g(A.indptr, A.indices, A.data)
    Z = np.stack([of.rademacher(4, of.probe_seed(1, p)) for p in range(8)], axis=1)
    loss, grad = of.hutchinson(A, G, Z)
    assert loss == pytest.approx(0.0, abs=1e-30)
    assert np.abs(grad.data).max() < 1e-15


def test_hutchinson_identity_probes_give_frobenius_norm():
    rp, ci, v, _ = cls.gen_aniso2d(5, 4, corr_len=2.0, sigma=0.5, ratio=20.0, theta=0.3, seed=2)
    n = 20
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    rng = np.random.default_rng(0)
    gi, gj, _ = of.tril_pattern(rp, ci)
    G = sp.csr_matrix((rng.uniform(0.1, 1.0, gi.size), (gi, gj)), shape=(n, n))
    E = (G.T @ A @ G).toarray() - np.eye(n)
    loss, _ = of.hutchinson(A, G, np.eye(n))
    assert loss == pytest.approx((E * E).sum() / (n * n), rel=1e-13)


def test_hutchinson_gradient_matches_finite_differences():
    rp, ci, v, _ = cls.gen_aniso2d(6, 5, corr_len=2.0, sigma=0.8, ratio=50.0, theta=0.9, seed=4)
    n = 30
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    rng = np.random.default_rng(1)
    gi, gj, _ = of.tril_pattern(rp, ci)
    gv = rng.uniform(-0.5, 0.5, gi.size) + (gi == gj)
    Z = np.stack([of.rademacher(n, of.probe_seed(3, p)) for p in range(8)], axis=1)

    def G_of(vals):
        return sp.csr_matrix((vals, (gi, gj)), shape=(n, n))

    _, grad = of.hutchinson(A, G_of(gv), Z)
    gdense = grad.toarray()
    h = 1e-5
    for k in rng.choice(gi.size, 25, replace=False):
        e = np.zeros_like(gv)
        e[k] = h
        fd = (of.hutchinson_loss(A, G_of(gv + e), Z) - of.hutchinson_loss(A, G_of(gv - e), Z)) / (2 * h)
        assert gdense[gi[k], gj[k]] == pytest.approx(fd, rel=1e-6, abs=1e-10)


def test_mlp_forward_zero_output_layer_is_bias():
    rng = np.random.default_rng(2)
    p = rng.normal(size=of.N_PARAMS).astype(np.float32)
    p[of.OFF_W3:of.OFF_B3] = 0
    p[of.OFF_B3] = 0.375
    X = rng.normal(size=(40, 16)).astype(np.float32)
    np.testing.assert_array_equal(of.mlp_forward(p, X), np.full(40, 0.375))


def test_mlp_backward_matches_finite_differences_fp64():
    rng = np.random.default_rng(3)
    p = rng.normal(scale=0.5, size=of.N_PARAMS)
    X = rng.normal(size=(37, 16))
    g = rng.normal(size=37)

    def f(pp):
        return float(g @ of.mlp_forward(pp, X, rounding=False))

    grad = of.mlp_backward(p, X, g, rounding=False)
    h = 1e-6
    idx = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False) for a, b in
                          [(of.OFF_W1, of.OFF_B1), (of.OFF_B1, of.OFF_W2), (of.OFF_W2, of.OFF_B2),
                           (of.OFF_B2, of.OFF_W3), (of.OFF_W3, of.OFF_B3)]] + [[of.OFF_B3]])
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (f(p + e) - f(p - e)) / (2 * h)
        assert grad[k] == pytest.approx(fd, rel=1e-5, abs=1e-7), k


def test_mlp_backward_rounding_close_to_exact():
    rng = np.random.default_rng(4)
    p = rng.normal(scale=0.3, size=of.N_PARAMS).astype(np.float32)
    X = of.bf16_round(rng.normal(size=(300, 16)))
    g = rng.normal(size=300)
    exact = of.mlp_backward(p.astype(np.float64), X.astype(np.float64), g, rounding=False)
    rounded = of.mlp_backward(p, X, g, rounding=True)
    for a, b in [(of.OFF_W1, of.OFF_B1), (of.OFF_W2, of.OFF_B2), (of.OFF_W3, of.OFF_B3)]:
        err = np.linalg.norm(exact[a:b] - rounded[a:b]) / np.linalg.norm(exact[a:b])
        assert err < 0.06  # ReLU masks flip where BF16 moves a pre-activation across 0


def _fp32_network(p, X, g, round_h1=True, round_d2=True):
    """The network in numpy FP32 (BLAS accumulation order, FP32 accumulators),
    the device's arithmetic restated independently of the oracle's FP64 one:
    (o, parameter gradient of sum g o)."""
    f = np.float32
    W1, b1, W2, b2, w3, b3 = (np.asarray(t, dtype=f) for t in of.split_params(p))
    W1q, W2q, Xq = of.bf16_round(W1), of.bf16_round(W2), of.bf16_round(X)
    a1 = (Xq @ W1q.T + b1).astype(f)
    h1 = of.bf16_round(np.maximum(a1, 0)) if round_h1 else np.maximum(a1, 0)
    a2 = (h1 @ W2q.T + b2).astype(f)
    h2 = of.bf16_round(np.maximum(a2, 0))
    o = (h2 @ w3).astype(f) + f(b3)
    g = np.asarray(g, dtype=f)
    d2 = np.where(a2 > 0, g[:, None] * w3[None, :], f(0)).astype(f)
    d2q = of.bf16_round(d2) if round_d2 else d2
    d1 = np.where(a1 > 0, (d2q @ W2q).astype(f), f(0)).astype(f)
    d1q = of.bf16_round(d1)
    grad = np.zeros(of.N_PARAMS)
    grad[of.OFF_W2:of.OFF_B2] = (d2q.T @ h1).ravel()
    grad[of.OFF_W1:of.OFF_B1] = (d1q.T @ Xq).ravel()
    grad[of.OFF_B1:of.OFF_W2] = d1.sum(axis=0, dtype=f)
    grad[of.OFF_B2:of.OFF_W3] = d2.sum(axis=0, dtype=f)
    grad[of.OFF_W3:of.OFF_B3] = (g[:, None] * h2).sum(axis=0, dtype=f)
    grad[of.OFF_B3] = g.sum(dtype=f)
    return o.astype(np.float64), grad


def test_mlp_bounds_hold_for_an_fp32_evaluation_and_catch_planted_errors():
    """The forward and backward parity bounds (what FP32 accumulation and the
    BF16 rounding flips it causes may change) contain an independent FP32
    evaluation of the network on real features, and are tight enough that
    leaving out one rounding step breaks them."""
    rng = np.random.default_rng(3)
    rp, ci, v, m = cls.gen_aniso2d(40, 36, corr_len=4.0, sigma=1.0, ratio=100.0, theta=0.6, seed=3)
    X = of.bf16_round(of.features(rp, ci, v))
    p = np.zeros(of.N_PARAMS, dtype=np.float32)
    p[:] = rng.normal(scale=0.3, size=of.N_PARAMS)
    g = rng.normal(scale=1e-3, size=X.shape[0]).astype(np.float32)
    o_ref, ob, fflips = of.mlp_forward_bound(p, X)
    np.testing.assert_array_equal(o_ref, of.mlp_forward(p, X))
    g_ref, gb, bflips = of.mlp_backward_bound(p, X, g)
    np.testing.assert_allclose(g_ref, of.mlp_backward(p, X, g), rtol=1e-12, atol=1e-300)
    assert 0 < fflips < X.shape[0] and bflips > 0
    o32, g32 = _fp32_network(p, X, g)
    assert np.all(np.abs(o32 - o_ref) <= ob)
    assert np.all(np.abs(g32 - g_ref) <= gb)
    # far tighter than a fixed 2 % of the output scale on most rows
    scale = np.abs(o_ref).max()
    assert np.median(ob) <= 1e-3 * scale and ob.max() <= 5e-2 * scale
    # planted: h1 left unrounded / d2 left unrounded
    o_bad, g_bad = _fp32_network(p, X, g, round_h1=False)
    assert np.any(np.abs(o_bad - o_ref) > ob)
    _, g_bad = _fp32_network(p, X, g, round_d2=False)
    assert np.any(np.abs(g_bad - g_ref) > gb)
