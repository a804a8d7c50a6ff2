"""CPU oracle of the learned factorized sparse approximate inverse.
TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker; the product never uses it.

The learned preconditioner has no counterpart in the reference
(src/library/solvers/preconditioners/ holds only the diagonal and void
preconditioners), so nothing here restates reference code. Each function
restates the arithmetic of clsparse_amd/csrc/fsai.hip in numpy:

* ``features``       -- per-nonzero features of G = tril(A)'s pattern;
* ``mlp_forward``    -- 16 -> 64 -> 64 -> 1 ReLU network with BF16 weights and
                        BF16 hidden activations, FP32 (here FP64) accumulation;
* ``mlp_backward``   -- its parameter gradient with the kernel's BF16 roundings;
* ``assemble_g``     -- G_ij = (delta_ij + o_ij) / sqrt(a_ii);
* ``rademacher``     -- the counter-based probe vectors;
* ``hutchinson``     -- L = ||(G^T A G - I) Z||_F^2 / (k n) and dL/dG;
* ``mlp_forward_bound`` / ``mlp_backward_bound`` -- the same network with a
                        per-output / per-parameter bound on what FP32
                        accumulation and the BF16 rounding flips it can cause
                        may change (the GPU parity bars).

How each is pinned (tests/test_oracle_fsai.py): ``hutchinson``'s gradient
against central finite differences of its own loss (a quadratic in G, so the
difference quotient is exact up to rounding) and, for a diagonal A with the
Jacobi G, against the closed form L = 0, dL/dG = 0; ``mlp_backward`` with
rounding switched off against finite differences of ``mlp_forward`` in FP64;
``features`` and ``assemble_g`` against hand-computed values on a 3x3
matrix; ``rademacher`` against values computed by hand from the SplitMix64
finaliser. ``mlp_forward`` with rounding is pinned only through
``mlp_backward``'s checks and the closed form o = b3 for w3 = 0. The two
bounds are pinned by an independent FP32 evaluation of the network in numpy
(other summation order, FP32 accumulators) that must stay inside them, and by
a planted error (a gate or a rounding step left out) that must not.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

N_FEAT = 16
N_HID = 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_HID * N_FEAT
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID
N_PARAMS = OFF_B3 + 1

_M64 = (1 << 64) - 1


def bf16_round(x) -> np.ndarray:
    """FP32 -> BF16 -> FP32, round to nearest even (finite inputs)."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(f.shape)


def bf16_bits_to_f32(bits: np.ndarray) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32)


def split_params(p: np.ndarray):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(N_HID, N_FEAT), p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(N_HID, N_HID), p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3],
            p[OFF_B3])


def tril_pattern(rp, ci):
    """(row, col, source index in A) of every nonzero of tril(A), row-major."""
    rp = np.asarray(rp, dtype=np.int64)
    ci = np.asarray(ci, dtype=np.int64)
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    keep = ci <= rows
    src = np.nonzero(keep)[0]
    return rows[keep], ci[keep], src


def dhat(rp, ci, v) -> np.ndarray:
    """1/sqrt(a_ii) (1 where a_ii is absent or not positive)."""
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(np.asarray(rp, dtype=np.int64)))
    d = np.ones(n)
    diag = np.asarray(ci) == rows
    # first stored diagonal of each row
    idx = np.nonzero(diag)[0]
    first = {}
    for k in idx[::-1]:
        first[rows[k]] = k
    for r, k in first.items():
        d[r] = v[k]
    out = np.ones(n)
    pos = d > 0
    out[pos] = 1.0 / np.sqrt(d[pos])
    return out


def features(rp, ci, v) -> np.ndarray:
    """FP32 features [nnz(tril A)][16]: scaled value a = a_ij dhat_i dhat_j,
    [a, i == j, r_i, r_j, s_i, s_j, a^2, len_i / 16, 0 x 8] with
    r_i = sum_{k != i} |a_ik| dhat_i dhat_k and s_i the signed sum."""
    rp = np.asarray(rp, dtype=np.int64)
    ci = np.asarray(ci, dtype=np.int64)
    v = np.asarray(v, dtype=np.float64)
    n = rp.size - 1
    dh = dhat(rp, ci, v)
    rows = np.repeat(np.arange(n), np.diff(rp))
    sc = v * dh[rows] * dh[ci]
    off = ci != rows
    r = np.zeros(n)
    s = np.zeros(n)
    np.add.at(r, rows[off], np.abs(sc[off]))
    np.add.at(s, rows[off], sc[off])
    r32, s32 = r.astype(np.float32), s.astype(np.float32)
    ln = (np.diff(rp).astype(np.float32) * np.float32(0.0625))
    gi, gj, src = tril_pattern(rp, ci)
    a = sc[src].astype(np.float32)
    X = np.zeros((gi.size, N_FEAT), dtype=np.float32)
    X[:, 0] = a
    X[:, 1] = (gi == gj).astype(np.float32)
    X[:, 2] = r32[gi]
    X[:, 3] = r32[gj]
    X[:, 4] = s32[gi]
    X[:, 5] = s32[gj]
    X[:, 6] = a * a
    X[:, 7] = ln[gi]
    return X


def mlp_forward(p, X, rounding: bool = True, return_cache: bool = False):
    """o = w3 . q(relu(W2 q(relu(W1 x + b1)) + b2)) + b3 per row of X, with q
    the BF16 rounding of the hidden activations and BF16 W1, W2 when
    ``rounding``; accumulation in FP64."""
    W1, b1, W2, b2, w3, b3 = split_params(p)
    if rounding:
        W1q = bf16_round(W1).astype(np.float64)
        W2q = bf16_round(W2).astype(np.float64)
        Xq = bf16_round(X).astype(np.float64)
    else:
        W1q, W2q, Xq = W1.astype(np.float64), W2.astype(np.float64), np.asarray(X, np.float64)
    a1 = Xq @ W1q.T + b1.astype(np.float64)
    h1 = np.maximum(a1, 0.0)
    if rounding:
        h1 = bf16_round(h1).astype(np.float64)
    a2 = h1 @ W2q.T + b2.astype(np.float64)
    h2 = np.maximum(a2, 0.0)
    if rounding:
        h2 = bf16_round(h2).astype(np.float64)
    o = h2 @ w3.astype(np.float64) + float(b3)
    if return_cache:
        return o, (Xq, W2q, a1, h1, a2, h2)
    return o


def mlp_backward(p, X, gout, rounding: bool = True) -> np.ndarray:
    """Parameter gradient of sum_k gout_k o_k (layout of the parameter vector).
    With ``rounding`` the backward products use BF16 copies of the deltas
    and of W2, as the MFMA kernels do; bias sums use the unrounded deltas."""
    _, _, _, _, w3, _ = split_params(p)
    o, (Xq, W2q, a1, h1, a2, h2) = mlp_forward(p, X, rounding, return_cache=True)
    g = np.asarray(gout, dtype=np.float64)
    q = (lambda t: bf16_round(t).astype(np.float64)) if rounding else (lambda t: t)
    d2 = (a2 > 0) * (g[:, None] * w3.astype(np.float64)[None, :])
    d2q = q(d2)
    d1 = (a1 > 0) * (d2q @ W2q)
    grad = np.zeros(N_PARAMS)
    grad[OFF_W2:OFF_B2] = (d2q.T @ h1).ravel()
    grad[OFF_W1:OFF_B1] = (q(d1).T @ Xq).ravel()
    grad[OFF_B1:OFF_W2] = d1.sum(axis=0)
    grad[OFF_B2:OFF_W3] = d2.sum(axis=0)
    grad[OFF_W3:OFF_B3] = (g[:, None] * h2).sum(axis=0)
    grad[OFF_B3] = g.sum()
    return grad


_U32 = 2.0 ** -24


def _relu_bf16(a) -> np.ndarray:
    """h = BF16(max(a, 0)) elementwise, as float64 (monotone in a)."""
    return bf16_round(np.maximum(a, 0.0)).astype(np.float64)


def _spread(f, lo, hi, mid):
    """max |f(t) - f(mid)| over t in [lo, hi] for a monotone f: the ends bound it."""
    fm = f(mid)
    return np.maximum(np.abs(f(hi) - fm), np.abs(f(lo) - fm))


def mlp_forward_bound(p, X):
    """(o, bound, flips): the oracle's BF16 forward output and, per row, a bound
    on |o_device - o| for any evaluation that accumulates each layer in FP32
    (in any order) where this one accumulates in FP64. A hidden activation is
    rounded to BF16 (after the ReLU) on both sides; it can land on another BF16
    value only when its pre-activation interval [a - e, a + e] (e: FP32
    accumulation bound of the layer, plus what flips of the layer below can
    move it) straddles a BF16 rounding boundary or zero. Those flips are
    counted (`flips`: rows with at least one) and their largest effect carried
    through the next layers; everything else differs only by FP32 rounding."""
    W1, b1, W2, b2, w3, b3 = split_params(p)
    W1q = bf16_round(W1).astype(np.float64)
    W2q = bf16_round(W2).astype(np.float64)
    Xq = bf16_round(X).astype(np.float64)
    a1 = Xq @ W1q.T + b1
    e1 = 2 * 17 * _U32 * (np.abs(Xq) @ np.abs(W1q).T + np.abs(b1))
    d1 = _spread(_relu_bf16, a1 - e1, a1 + e1, a1)
    h1 = _relu_bf16(a1)
    a2 = h1 @ W2q.T + b2
    E2 = d1 @ np.abs(W2q).T + 2 * 65 * _U32 * ((h1 + d1) @ np.abs(W2q).T + np.abs(b2))
    d2 = _spread(_relu_bf16, a2 - E2, a2 + E2, a2)
    h2 = _relu_bf16(a2)
    o = h2 @ w3 + float(b3)
    bound = d2 @ np.abs(w3) + 2 * 66 * _U32 * ((h2 + d2) @ np.abs(w3) + abs(float(b3)))
    flips = int(np.count_nonzero((d1 > 0).any(axis=1) | (d2 > 0).any(axis=1)))
    return o, bound, flips


def mlp_backward_bound(p, X, gout):
    """(grad, bound, flips): ``mlp_backward`` with rounding and a per-parameter
    bound on |grad_device - grad| for an evaluation that accumulates in FP32
    over all nonzeros (any order) and may round each BF16 value (hidden
    activations, deltas) to the neighbouring BF16 value wherever the FP32 /
    FP64 difference of its argument, and the flips below it, allow. gout may
    differ by one FP32 rounding. `flips`: nonzeros with any possible flip."""
    W1, b1, W2, b2, w3, b3 = split_params(p)
    W1q = bf16_round(W1).astype(np.float64)
    W2q = bf16_round(W2).astype(np.float64)
    Xq = bf16_round(X).astype(np.float64)
    g = np.asarray(gout, dtype=np.float64)
    N = g.size
    gam = 2 * max(N, 1) * _U32
    q = lambda t: bf16_round(t).astype(np.float64)
    # forward, with the activation deviations of mlp_forward_bound
    a1 = Xq @ W1q.T + b1
    e1 = 2 * 17 * _U32 * (np.abs(Xq) @ np.abs(W1q).T + np.abs(b1))
    dh1 = _spread(_relu_bf16, a1 - e1, a1 + e1, a1)
    h1 = _relu_bf16(a1)
    a2 = h1 @ W2q.T + b2
    E2 = dh1 @ np.abs(W2q).T + 2 * 65 * _U32 * ((h1 + dh1) @ np.abs(W2q).T + np.abs(b2))
    dh2 = _spread(_relu_bf16, a2 - E2, a2 + E2, a2)
    h2 = _relu_bf16(a2)
    dg = 2 * _U32 * np.abs(g)
    # D2 = relu'(a2) g w3 (FP32 product on the device), then BF16
    v2 = g[:, None] * w3[None, :]
    ev2 = _U32 * np.abs(v2) + dg[:, None] * np.abs(w3)[None, :]
    on2 = a2 > 0
    unc2 = (a2 - E2 <= 0) & (a2 + E2 > 0)
    d2 = np.where(on2, v2, 0.0)
    lo2 = np.where(unc2, np.minimum(0.0, v2 - ev2), np.where(on2, v2 - ev2, 0.0))
    hi2 = np.where(unc2, np.maximum(0.0, v2 + ev2), np.where(on2, v2 + ev2, 0.0))
    d2q = q(d2)
    dd2q = _spread(q, lo2, hi2, d2)
    dd2 = np.maximum(np.abs(hi2 - d2), np.abs(lo2 - d2))
    # D1 = relu'(a1) (D2q W2q) (FP32 MFMA), then BF16
    c1 = d2q @ W2q
    ec1 = dd2q @ np.abs(W2q) + 2 * 65 * _U32 * ((np.abs(d2q) + dd2q) @ np.abs(W2q))
    on1 = a1 > 0
    unc1 = (a1 - e1 <= 0) & (a1 + e1 > 0)
    d1 = np.where(on1, c1, 0.0)
    lo1 = np.where(unc1, np.minimum(0.0, c1 - ec1), np.where(on1, c1 - ec1, 0.0))
    hi1 = np.where(unc1, np.maximum(0.0, c1 + ec1), np.where(on1, c1 + ec1, 0.0))
    d1q = q(d1)
    dd1q = _spread(q, lo1, hi1, d1)
    dd1 = np.maximum(np.abs(hi1 - d1), np.abs(lo1 - d1))
    grad = np.zeros(N_PARAMS)
    bound = np.zeros(N_PARAMS)
    grad[OFF_W2:OFF_B2] = (d2q.T @ h1).ravel()
    bound[OFF_W2:OFF_B2] = (dd2q.T @ (h1 + dh1) + np.abs(d2q).T @ dh1 +
                            gam * ((np.abs(d2q) + dd2q).T @ (h1 + dh1))).ravel()
    grad[OFF_W1:OFF_B1] = (d1q.T @ Xq).ravel()
    bound[OFF_W1:OFF_B1] = (dd1q.T @ np.abs(Xq) + gam * ((np.abs(d1q) + dd1q).T @ np.abs(Xq))).ravel()
    grad[OFF_B1:OFF_W2] = d1.sum(axis=0)
    bound[OFF_B1:OFF_W2] = dd1.sum(axis=0) + gam * (np.abs(d1) + dd1).sum(axis=0)
    grad[OFF_B2:OFF_W3] = d2.sum(axis=0)
    bound[OFF_B2:OFF_W3] = dd2.sum(axis=0) + gam * (np.abs(d2) + dd2).sum(axis=0)
    grad[OFF_W3:OFF_B3] = (g[:, None] * h2).sum(axis=0)
    bound[OFF_W3:OFF_B3] = (np.abs(g)[:, None] * dh2 + dg[:, None] * (h2 + dh2)).sum(axis=0) + \
        gam * (np.abs(g)[:, None] * (h2 + dh2)).sum(axis=0)
    grad[OFF_B3] = g.sum()
    bound[OFF_B3] = dg.sum() + gam * np.abs(g).sum()
    flips = int(np.count_nonzero((dh1 > 0).any(axis=1) | (dh2 > 0).any(axis=1) | unc2.any(axis=1) |
                                 (dd2q > 0).any(axis=1) | (dd1q > 0).any(axis=1)))
    return grad, bound, flips


def assemble_g(rp, ci, v, o=None):
    """G (scipy CSR, lower triangular): G_ij = (delta_ij + o_ij) dhat_i."""
    rp = np.asarray(rp, dtype=np.int64)
    n = rp.size - 1
    gi, gj, _ = tril_pattern(rp, ci)
    dh = dhat(rp, ci, v)
    oo = np.zeros(gi.size) if o is None else np.asarray(o, dtype=np.float64)
    val = ((gi == gj).astype(np.float64) + oo) * dh[gi]
    return sp.csr_matrix((val, (gi, gj)), shape=(n, n))


def _mix(x):
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9 & _M64
    x = (x ^ (x >> 27)) * 0x94D049BB133111EB & _M64
    return x ^ (x >> 31)


def probe_seed(seed: int, probe: int) -> int:
    return (seed * 0x100000001B3 + probe) & _M64


def rademacher(n: int, seed: int) -> np.ndarray:
    """z_i = +1 if the SplitMix64 finaliser of seed ^ (golden * (i + 1)) is odd, else -1."""
    i = np.arange(n, dtype=np.uint64) + np.uint64(1)
    with np.errstate(over="ignore"):
        x = np.uint64(seed) ^ (np.uint64(0x9E3779B97F4A7C15) * i)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return np.where(x & np.uint64(1), 1.0, -1.0)


def hutchinson(A, G, Z):
    """Loss ||(G^T A G - I) Z||_F^2 / (k n) and its gradient on G's pattern
    (as a CSR matrix with G's structure): dL/dG_ij = 2/(k n) sum_z
    (u_i v_j + w_i z_j), t = G z, u = A t, v = G^T u - z, w = A G v."""
    n, k = Z.shape
    G = sp.csr_matrix(G)
    U = A @ (G @ Z)
    V = G.T @ U - Z
    W = A @ (G @ V)
    loss = float((V * V).sum()) / (k * n)
    coo = G.tocoo()
    gval = 2.0 / (k * n) * (np.einsum("ij,ij->i", U[coo.row], V[coo.col]) +
                            np.einsum("ij,ij->i", W[coo.row], Z[coo.col]))
    grad = sp.csr_matrix((gval, (coo.row, coo.col)), shape=G.shape)
    return loss, grad


def hutchinson_loss(A, G, Z) -> float:
    n, k = Z.shape
    E = G.T @ (A @ (G @ Z)) - Z
    return float((E * E).sum()) / (k * n)
