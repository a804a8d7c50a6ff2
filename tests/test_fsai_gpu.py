"""Learned FSAI on the GPU against the CPU oracle (oracle/fsai.py).

Features, G assembly and the Hutchinson gradient are compared to the oracle on
the same inputs; the MFMA forward and backward kernels to the oracle's
BF16-emulating network; the solver path to its own Jacobi fallback; and a
short training run shows what the network gains over the method without it
(zero network = Jacobi).
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import clsparse_amd as cls
from oracle import fsai as of

from gpu_helpers import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctl():
    c = cls.Control()
    yield c
    c.release()


def _aniso(nx, ny, seed, ratio=100.0, theta=0.6, sigma=1.0, corr=4.0):
    rp, ci, v, m = cls.gen_aniso2d(nx, ny, corr_len=corr, sigma=sigma, ratio=ratio, theta=theta,
                                   seed=seed)
    return rp, ci, v, m


def _random_params(seed, scale=0.2):
    rng = np.random.default_rng(seed)
    p = np.zeros(of.N_PARAMS, dtype=np.float32)
    p[of.OFF_W1:of.OFF_B1] = rng.normal(scale=0.5, size=of.OFF_B1 - of.OFF_W1)
    p[of.OFF_B1:of.OFF_W2] = rng.normal(scale=0.1, size=of.N_HID)
    p[of.OFF_W2:of.OFF_B2] = rng.normal(scale=0.18, size=of.N_HID * of.N_HID)
    p[of.OFF_B2:of.OFF_W3] = rng.normal(scale=0.1, size=of.N_HID)
    p[of.OFF_W3:of.OFF_B3] = rng.normal(scale=scale / 8, size=of.N_HID)
    p[of.OFF_B3] = 0.01
    return p


def _sample_matrix(ctl, nx=24, ny=20, seed=3):
    rp, ci, v, m = _aniso(nx, ny, seed)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    return rp, ci, v, m, A


def test_features_and_jacobi_factor_match_oracle(ctl):
    rp, ci, v, m, A = _sample_matrix(ctl)
    say("fsai: features 24x20 aniso")
    f = cls.Fsai(A, ctl)
    try:
        X = of.bf16_bits_to_f32(f.features()).reshape(-1, 16)
        Xo = of.bf16_round(of.features(rp, ci, v))
        assert X.shape == Xo.shape
        # identical up to one BF16 step where FP64 summation order moves a row sum
        diff = np.abs(X - Xo)
        assert np.all(diff <= 2.0 ** -7 * np.maximum(np.abs(Xo), 1e-30) + 1e-30)
        assert (diff > 0).mean() < 0.01
        np.testing.assert_array_equal(X[:, 0:2], Xo[:, 0:2])
        G = of.assemble_g(rp, ci, v)
        # a new factor is Jacobi's, assembled where it is first read
        np.testing.assert_array_equal(f.values()[2], G.data)
        f.apply_net(None, ctl)
        grp, gci, gv = f.values()
        np.testing.assert_array_equal(grp, G.indptr)
        np.testing.assert_array_equal(gci, G.indices)
        np.testing.assert_array_equal(gv, G.data)  # 1/sqrt(a_ii): correctly rounded both sides
    finally:
        f.release()


def test_update_values_gives_the_new_matrix_its_jacobi_factor(ctl):
    """clsparseAmdFsaiUpdateValues on new values of the same pattern: the
    features and G = D^-1/2 follow the new matrix (G assembled where it is
    first read), and a solve with the factor converges on it."""
    rp, ci, v, m, A = _sample_matrix(ctl, 30, 26, seed=12)
    f = cls.Fsai(A, ctl)
    try:
        v2 = v * 3.0
        A2 = cls.CsrMatrix(rp, ci, v2, m)
        A2.meta_create(ctl)
        say("fsai: update values")
        f.update_values(A2, ctl)
        np.testing.assert_array_equal(f.values()[2], of.assemble_g(rp, ci, v2).data)
        b = cls.gen_uniform(m, -1, 1, seed=3)
        st, x = _solve_fsai(ctl, A2, f, b)
        Acpu = sp.csr_matrix((v2, ci, rp), shape=(m, m))
        assert st.status_detail == 0
        assert np.linalg.norm(b - Acpu @ x) / np.linalg.norm(b) <= 1e-8
    finally:
        f.release()


def test_mlp_forward_mfma_matches_oracle(ctl):
    rp, ci, v, m, A = _sample_matrix(ctl, 33, 17, seed=5)  # nnz not a multiple of 16
    net = cls.Mlp(seed=7)
    f = cls.Fsai(A, ctl)
    try:
        p = _random_params(11)
        net.set_params(p)
        say("fsai: MLP forward vs oracle")
        f.apply_net(net, ctl)
        o = f.output().astype(np.float64)
        X = of.bf16_bits_to_f32(f.features()).reshape(-1, 16)
        # Per nonzero: within what FP32 accumulation and the BF16 rounding
        # flips it can cause allow (flips counted by the oracle).
        o_ref, bound, flips = of.mlp_forward_bound(p, X)
        err = np.abs(o - o_ref)
        bad = err > bound
        say(f"  forward: {flips} of {o.size} nonzeros with a possible BF16 flip, "
            f"max err / bound {float((err / np.maximum(bound, 1e-300)).max()):.3f}")
        assert not bad.any(), (int(np.argmax(bad)), float(err[bad][0]), float(bound[bad][0]))
        # G assembled from the network output
        _, _, gv = f.values()
        G = of.assemble_g(rp, ci, v, o)
        np.testing.assert_allclose(gv, G.data, rtol=1e-15, atol=0)
    finally:
        f.release()
        net.release()


def test_new_network_reproduces_jacobi(ctl):
    rp, ci, v, m, A = _sample_matrix(ctl)
    net = cls.Mlp(seed=3)
    f = cls.Fsai(A, ctl)
    try:
        say("fsai: new network = Jacobi")
        f.apply_net(net, ctl)
        assert np.all(f.output() == 0)
        _, _, gv = f.values()
        np.testing.assert_array_equal(gv, of.assemble_g(rp, ci, v).data)
    finally:
        f.release()
        net.release()


def test_hutchinson_loss_gradient_and_backward_match_oracle(ctl):
    rp, ci, v, m, A = _sample_matrix(ctl, 20, 18, seed=8)
    n = m
    net = cls.Mlp(seed=2)
    f = cls.Fsai(A, ctl)
    try:
        p = _random_params(5)
        net.set_params(p)
        say("fsai: Hutchinson gradient + MLP backward vs oracle")
        probes, seed = 8, 77
        loss, gG = f.train_step(net, A, ctl, probes=probes, seed=seed, lr=0.0, want_grad=True)
        o = f.output().astype(np.float64)
        X = of.bf16_bits_to_f32(f.features()).reshape(-1, 16)
        Acpu = sp.csr_matrix((v, ci, rp), shape=(n, n))
        G = of.assemble_g(rp, ci, v, o)
        Z = np.stack([of.rademacher(n, of.probe_seed(seed, k)) for k in range(probes)], axis=1)
        loss_ref, grad_ref = of.hutchinson(Acpu, G, Z)
        assert loss == pytest.approx(loss_ref, rel=1e-11)
        gi, gj, _ = of.tril_pattern(rp, ci)
        gref = np.asarray(grad_ref[gi, gj]).ravel()
        np.testing.assert_allclose(gG, gref, rtol=1e-9, atol=1e-12 * np.abs(gref).max())
        # parameter gradient through the network: gout = dL/dG * dhat_i
        gout = gG * of.dhat(rp, ci, v)[gi]
        g_ref, bound, flips = of.mlp_backward_bound(p, X, gout.astype(np.float32))
        g_dev = net.grad().astype(np.float64)
        # Per parameter: within what FP32 accumulation over the nonzeros and
        # the BF16 rounding flips (activations, deltas) can change.
        err = np.abs(g_dev - g_ref)
        bad = err > bound
        say(f"  backward: {flips} nonzeros with a possible flip; max err / bound "
            f"{float((err / np.maximum(bound, 1e-300)).max()):.3f}")
        assert not bad.any(), (int(np.argmax(bad)), float(err[bad][0]), float(bound[bad][0]))
    finally:
        f.release()
        net.release()


def _solve_fsai(ctl, A, f, b, tol=1e-8, fp64_inner=False, max_iters=20000):
    lib = cls.load()
    n = A.num_rows
    xd = cls.DenseVector(torch.zeros(n, dtype=torch.float64, device="cuda"))
    bd = cls.DenseVector(torch.tensor(b, dtype=torch.float64, device="cuda"))
    sc = cls.SolverControl(cls.DIAGONAL, max_iters, tol, 0.0)
    try:
        fn = lib.clsparseAmdDcsrcgFsai64 if fp64_inner else lib.clsparseAmdDcsrcgFsai
        cls.check("clsparseAmdDcsrcgFsai", fn(ctypes.byref(xd.s), ctypes.byref(A.s),
                                               ctypes.byref(bd.s), f.handle, sc.handle,
                                               ctl.handle))
        return sc.stats(), xd.t.cpu().numpy()
    finally:
        sc.release()


def _solve_jacobi_mixed(ctl, A, b, tol=1e-8, max_iters=20000):
    lib = cls.load()
    n = A.num_rows
    xd = cls.DenseVector(torch.zeros(n, dtype=torch.float64, device="cuda"))
    bd = cls.DenseVector(torch.tensor(b, dtype=torch.float64, device="cuda"))
    sc = cls.SolverControl(cls.DIAGONAL, max_iters, tol, 0.0)
    try:
        cls.check("clsparseAmdDcsrcgMixed", lib.clsparseAmdDcsrcgMixed(
            ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s), sc.handle, ctl.handle))
        return sc.stats(), xd.t.cpu().numpy()
    finally:
        sc.release()


def test_fsai_solver_with_jacobi_factor_matches_jacobi(ctl, monkeypatch):
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_POLICY", "always")
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_BF16", "0")  # FP32 G = D^-1/2 exactly
    rp, ci, v, m, A = _sample_matrix(ctl, 40, 40, seed=12)
    b = cls.gen_uniform(m, -1, 1, seed=4)
    f = cls.Fsai(A, ctl)
    try:
        say("fsai: PCG with G = D^-1/2 vs Jacobi PCG")
        st_f, x_f = _solve_fsai(ctl, A, f, b)
        st_j, x_j = _solve_jacobi_mixed(ctl, A, b)
        assert st_f.status_detail == 0 and st_j.status_detail == 0
        assert st_f.preconditioner == 2 and st_j.preconditioner == 1
        assert abs(st_f.iterations - st_j.iterations) <= max(3, st_j.iterations // 50)
        Acpu = sp.csr_matrix((v, ci, rp), shape=(m, m))
        assert np.linalg.norm(b - Acpu @ x_f) / np.linalg.norm(b) <= 1e-8
    finally:
        f.release()


def test_non_finite_factor_falls_back_to_jacobi(ctl):
    rp, ci, v, m, A = _sample_matrix(ctl, 30, 30, seed=13)
    b = cls.gen_uniform(m, -1, 1, seed=5)
    net = cls.Mlp(seed=1)
    f = cls.Fsai(A, ctl)
    try:
        p = _random_params(3)
        p[of.OFF_B3] = np.inf
        net.set_params(p)
        say("fsai: non-finite factor -> Jacobi")
        f.apply_net(net, ctl)
        st, x = _solve_fsai(ctl, A, f, b)
        assert st.preconditioner == 3 and st.status_detail == 0
        Acpu = sp.csr_matrix((v, ci, rp), shape=(m, m))
        assert np.linalg.norm(b - Acpu @ x) / np.linalg.norm(b) <= 1e-8
    finally:
        f.release()
        net.release()


def test_training_lowers_loss_and_beats_the_untrained_method(ctl, monkeypatch):
    """Trains on 6 seeded matrices, evaluates on a held-out one: the learned G
    lowers the Hutchinson loss and the PCG iteration count below the same
    solver with a zero network (= Jacobi)."""
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_POLICY", "always")
    net = cls.Mlp(seed=21)
    train = []
    for s in range(6):
        rp, ci, v, m = _aniso(48, 48, seed=100 + s, ratio=float(10 ** (1 + s % 3)),
                              theta=0.5 * s)
        A = cls.CsrMatrix(rp, ci, v, m)
        A.meta_create(ctl)
        train.append((A, cls.Fsai(A, ctl)))
    rp, ci, v, m = _aniso(48, 48, seed=999, ratio=100.0, theta=1.1)
    Ah = cls.CsrMatrix(rp, ci, v, m)
    Ah.meta_create(ctl)
    fh = cls.Fsai(Ah, ctl)
    try:
        say("fsai: training 6 matrices x 40 epochs")
        loss0, _ = fh.train_step(net, Ah, ctl, probes=8, seed=5, lr=0.0)
        for ep in range(40):
            for k, (A, f) in enumerate(train):
                f.train_step(net, A, ctl, probes=8, seed=1000 * ep + k, lr=2e-3)
        loss1, _ = fh.train_step(net, Ah, ctl, probes=8, seed=5, lr=0.0)
        say(f"fsai: held-out loss {loss0:.4f} -> {loss1:.4f}")
        assert loss1 < 0.8 * loss0
        b = cls.gen_uniform(m, -1, 1, seed=6)
        fh.apply_net(net, ctl)
        st_l, x_l = _solve_fsai(ctl, Ah, fh, b)
        fh.apply_net(None, ctl)
        st_0, _ = _solve_fsai(ctl, Ah, fh, b)
        say(f"fsai: held-out iterations learned {st_l.iterations} vs zero network "
            f"{st_0.iterations}")
        assert st_l.status_detail == 0 and st_l.preconditioner == 2
        assert st_l.iterations < 0.9 * st_0.iterations
        Acpu = sp.csr_matrix((v, ci, rp), shape=(m, m))
        assert np.linalg.norm(b - Acpu @ x_l) / np.linalg.norm(b) <= 1e-8
    finally:
        for _, f in train:
            f.release()
        fh.release()
        net.release()


@pytest.mark.parametrize("dia", ["1", "0"])
def test_solve_after_a_training_step_uses_its_network_output(ctl, monkeypatch, dia):
    """A training step assembles only the FP64 factor; the next solve brings
    the FP32 / BF16 / diagonal copies up to date from the step's network
    output, so it runs exactly as after apply_net with the same network."""
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_POLICY", "always")
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_DIA", dia)
    rp, ci, v, m, A = _sample_matrix(ctl, 24, 24, seed=31)
    b = cls.gen_uniform(m, -1, 1, seed=32)
    net = cls.Mlp(seed=9)
    f1, f2 = cls.Fsai(A, ctl), cls.Fsai(A, ctl)
    try:
        net.set_params(_random_params(33, scale=0.05))
        say(f"fsai: solve after a training step (diagonal storage {dia})")
        f1.train_step(net, A, ctl, probes=4, seed=7, lr=0.0)  # lr 0: the network stays
        st1, x1 = _solve_fsai(ctl, A, f1, b)
        f2.apply_net(net, ctl)
        st2, x2 = _solve_fsai(ctl, A, f2, b)
        assert st1.status_detail == 0 and st1.preconditioner == 2
        assert st1.iterations == st2.iterations
        np.testing.assert_array_equal(x1, x2)
        np.testing.assert_array_equal(f1.values()[2], f2.values()[2])
    finally:
        f1.release()
        f2.release()
        net.release()


def test_pilot_race_picks_a_preconditioner_and_converges(ctl, monkeypatch):
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_POLICY", "race")
    rp, ci, v, m, A = _sample_matrix(ctl, 64, 64, seed=14)
    b = cls.gen_uniform(m, -1, 1, seed=7)
    f = cls.Fsai(A, ctl)
    try:
        say("fsai: pilot race")
        st, x = _solve_fsai(ctl, A, f, b)
        assert st.status_detail == 0 and st.preconditioner in (2, 3)
        raced = st.iterations > 3 * 32
        if raced:
            assert st.pilot_rate_fsai > 0 and st.pilot_rate_jacobi > 0
            # Jacobi must beat the learned factor by the race margin (1.15)
            assert (st.preconditioner == 2) == (st.pilot_rate_jacobi <= 1.15 * st.pilot_rate_fsai)
        Acpu = sp.csr_matrix((v, ci, rp), shape=(m, m))
        assert np.linalg.norm(b - Acpu @ x) / np.linalg.norm(b) <= 1e-8
        # a second solve with the same factor reuses the outcome without racing
        st2, x2 = _solve_fsai(ctl, A, f, b)
        assert st2.status_detail == 0 and st2.preconditioner == st.preconditioner
        if raced:
            assert (st2.pilot_rate_fsai, st2.pilot_rate_jacobi) == (st.pilot_rate_fsai,
                                                                    st.pilot_rate_jacobi)
            assert st2.iterations <= st.iterations
        assert np.linalg.norm(b - Acpu @ x2) / np.linalg.norm(b) <= 1e-8
        # a changed factor forgets it
        f.apply_net(None, ctl)
        st3, _ = _solve_fsai(ctl, A, f, b)
        if st3.iterations > 3 * 32:
            assert st3.pilot_rate_jacobi > 0 and st3.pilot_rate_jacobi != st.pilot_rate_jacobi
    finally:
        f.release()


def test_loss_estimate_decides_whether_to_race(ctl, monkeypatch):
    """By default a new factor in diagonal storage is first judged by
    Hutchinson estimates of ||F^T A F - I||_F^2 / n (one probe, the factor
    as the solver applies it, and Jacobi's D^-1/2): the trained network's
    factor, well under half of Jacobi's, is kept without the pilot race; the
    zero network's (G = D^-1/2) races. The estimates match the training
    step's FP64 Hutchinson loss over eight other probes."""
    monkeypatch.delenv("CLSPARSE_AMD_FSAI_POLICY", raising=False)
    from clsparse_amd import fsai as fs
    params, _ = fs.load_weights()
    rp, ci, v, m = cls.gen_diffusion3d(48, 44, 40, sigma=1.0, ax=1.0, ay=1.0, az=100.0, seed=9)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    b = cls.gen_uniform(m, -1, 1, seed=19)
    Acpu = sp.csr_matrix((v, ci, rp), shape=(m, m))
    net, net0 = cls.Mlp(seed=1), cls.Mlp(seed=1)
    net.set_params(params)
    f = cls.Fsai(A, ctl)
    try:
        loss_l = f.train_step(net, A, ctl, probes=8, seed=5, lr=0.0)[0]
        loss_0 = f.train_step(net0, A, ctl, probes=8, seed=5, lr=0.0)[0]
        f.apply_net(net, ctl)
        assert len(f.diagonals()) == 4
        say("fsai: loss estimate, trained network")
        st, x = _solve_fsai(ctl, A, f, b)
        assert st.loss_fsai == pytest.approx(loss_l, rel=0.05), (st.loss_fsai, loss_l)
        assert st.loss_jacobi == pytest.approx(loss_0, rel=0.05), (st.loss_jacobi, loss_0)
        assert st.loss_fsai <= 0.5 * st.loss_jacobi
        assert st.preconditioner == 2 and st.pilot_rate_fsai == 0 and st.pilot_rate_jacobi == 0
        assert st.status_detail == 0 and np.linalg.norm(b - Acpu @ x) / np.linalg.norm(b) <= 1e-8
        # the outcome stays with the factor: no estimate on the next solve
        st2, _ = _solve_fsai(ctl, A, f, b)
        assert st2.loss_fsai == 0 and st2.preconditioner == 2 and st2.iterations <= st.iterations
        f.apply_net(None, ctl)  # G = D^-1/2: no lead over Jacobi, so the race runs
        say("fsai: loss estimate, zero network")
        st3, x3 = _solve_fsai(ctl, A, f, b)
        assert st3.loss_fsai == pytest.approx(st3.loss_jacobi, rel=0.05)
        assert st3.pilot_rate_fsai > 0 and st3.pilot_rate_jacobi > 0
        assert st3.status_detail == 0 and np.linalg.norm(b - Acpu @ x3) / np.linalg.norm(b) <= 1e-8
        # a new factor forgets the race it ran: judged afresh, no rates left over
        f.apply_net(net, ctl)
        st5, _ = _solve_fsai(ctl, A, f, b)
        assert st5.loss_fsai > 0 and st5.pilot_rate_fsai == 0 and st5.pilot_rate_jacobi == 0
        # CLSPARSE_AMD_FSAI_POLICY=race races whatever the estimate would say
        monkeypatch.setenv("CLSPARSE_AMD_FSAI_POLICY", "race")
        f.apply_net(net, ctl)
        st4, _ = _solve_fsai(ctl, A, f, b)
        assert st4.loss_fsai == 0 and st4.pilot_rate_fsai > 0 and st4.pilot_rate_jacobi > 0
    finally:
        f.release()
        net.release()
        net0.release()


@pytest.mark.parametrize("margin,expect", [("1e-6", 3), ("1e6", 2)])
def test_race_margin_decides_the_pilot_race(ctl, monkeypatch, margin, expect):
    """CLSPARSE_AMD_RACE_MARGIN scales Jacobi's required lead: a tiny margin
    hands any raced solve to Jacobi, a huge one keeps the learned factor;
    both converge."""
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_POLICY", "race")
    monkeypatch.setenv("CLSPARSE_AMD_RACE_MARGIN", margin)
    rp, ci, v, m, A = _sample_matrix(ctl, 64, 64, seed=15)
    b = cls.gen_uniform(m, -1, 1, seed=8)
    f = cls.Fsai(A, ctl)
    try:
        say(f"fsai: race margin {margin}")
        st, x = _solve_fsai(ctl, A, f, b)
        assert st.status_detail == 0
        if st.iterations > 3 * 32:  # the race ran to its end
            assert st.pilot_rate_fsai > 0 and st.pilot_rate_jacobi > 0
            assert st.preconditioner == expect
        Acpu = sp.csr_matrix((v, ci, rp), shape=(m, m))
        assert np.linalg.norm(b - Acpu @ x) / np.linalg.norm(b) <= 1e-8
    finally:
        f.release()


def test_bf16_factor_matches_fp32_factor(ctl, monkeypatch):
    """The BF16 copy of a trained G preconditions as well as the FP32 one."""
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_POLICY", "always")
    from clsparse_amd import fsai as fs
    params, _ = fs.load_weights()
    rp, ci, v, m = _aniso(64, 64, seed=77, ratio=200.0, theta=0.3)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    net = cls.Mlp(seed=1)
    net.set_params(params)
    f = cls.Fsai(A, ctl)
    try:
        f.apply_net(net, ctl)
        b = cls.gen_uniform(m, -1, 1, seed=9)
        say("fsai: BF16 vs FP32 factor")
        monkeypatch.setenv("CLSPARSE_AMD_FSAI_BF16", "0")
        st32, _ = _solve_fsai(ctl, A, f, b)
        monkeypatch.setenv("CLSPARSE_AMD_FSAI_BF16", "1")
        st16, x16 = _solve_fsai(ctl, A, f, b)
        assert st32.status_detail == 0 and st16.status_detail == 0
        assert st16.iterations <= 1.05 * st32.iterations + 8
        Acpu = sp.csr_matrix((v, ci, rp), shape=(m, m))
        assert np.linalg.norm(b - Acpu @ x16) / np.linalg.norm(b) <= 1e-8
    finally:
        f.release()
        net.release()


def test_missing_diagonal_factor_falls_back_to_jacobi(ctl):
    """A row of A without a stored diagonal makes G singular: the factor is
    marked invalid when it is assembled and the solver runs Jacobi from the
    start (that matrix is not positive definite, so only the choice of
    preconditioner and a clean return are checked)."""
    rp, ci, v, m = _aniso(20, 20, seed=31)
    a = sp.csr_matrix((v, ci, rp), shape=(m, m)).tolil()
    a[5, 5] = 0.0
    a = sp.csr_matrix(a)
    a.eliminate_zeros()
    a.sort_indices()
    assert a[5, 5] == 0 and a.indptr[6] - a.indptr[5] > 0
    A = cls.CsrMatrix(a.indptr.astype(np.uint32), a.indices.astype(np.uint32), a.data, m)
    A.meta_create(ctl)
    f = cls.Fsai(A, ctl)
    try:
        b = cls.gen_uniform(m, -1, 1, seed=3)
        say("fsai: missing diagonal -> Jacobi")
        st, x = _solve_fsai(ctl, A, f, b, max_iters=50)
        assert st.preconditioner == 3
        assert st.iterations <= 50
    finally:
        f.release()


def _tril_offsets(rp, ci):
    """Distinct column - row offsets of the lower triangle, ascending."""
    rows = np.repeat(np.arange(rp.size - 1, dtype=np.int64), np.diff(rp.astype(np.int64)))
    d = ci.astype(np.int64) - rows
    return np.unique(d[d <= 0]).tolist(), int((d <= 0).sum())


def test_factor_storage_follows_the_pattern(ctl):
    """Stencil factors are applied from diagonal storage; a pattern on many
    diagonals keeps the CSR SpMVs."""
    rp3, ci3, v3, m3 = cls.gen_diffusion3d(12, 10, 8, sigma=1.0, ax=1.0, ay=1.0, az=100.0, seed=3)
    rp2, ci2, v2, m2 = _aniso(24, 20, seed=3)
    a = sp.random(600, 600, density=0.02, random_state=np.random.default_rng(5), format="csr")
    a = a + a.T + sp.identity(600) * 50.0
    a = sp.csr_matrix(a)
    a.sort_indices()
    cases = [(rp3, ci3, v3, m3), (rp2, ci2, v2, m2),
             (a.indptr.astype(np.uint32), a.indices.astype(np.uint32), a.data, 600)]
    say("fsai: factor storage by pattern")
    for k, (rp, ci, v, m) in enumerate(cases):
        A = cls.CsrMatrix(rp, ci, v, m)
        A.meta_create(ctl)
        f = cls.Fsai(A, ctl)
        try:
            offs, nnz_tril = _tril_offsets(rp, ci)
            expect = offs if (len(offs) <= 8 and len(offs) * m <= 1.5 * nnz_tril) else []
            assert f.diagonals() == expect, (k, f.diagonals(), offs)
        finally:
            f.release()
    assert _tril_offsets(rp3, ci3)[0] == [-120, -12, -1, 0]
    assert len(_tril_offsets(rp2, ci2)[0]) == 5


@pytest.mark.parametrize("dia", ["1", "0"])
def test_first_iteration_applies_the_bf16_factor(ctl, monkeypatch, dia):
    """One PCG iteration from x = 0 gives x = alpha G (G^T b): its direction
    checks the factor application (diagonal storage, or the CSR SpMVs of G
    and G^T) against the product with the BF16-rounded factor."""
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_POLICY", "always")
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_DIA", dia)
    from clsparse_amd import fsai as fs
    params, _ = fs.load_weights()
    rp, ci, v, m = cls.gen_diffusion3d(16, 14, 12, sigma=1.0, ax=1.0, ay=1.0, az=100.0, seed=4)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    net = cls.Mlp(seed=1)
    net.set_params(params)
    f = cls.Fsai(A, ctl)
    try:
        f.apply_net(net, ctl)
        assert len(f.diagonals()) == 4
        b = cls.gen_uniform(m, -1, 1, seed=12)
        say(f"fsai: first iteration, diagonal storage={dia}")
        st, x = _solve_fsai(ctl, A, f, b, max_iters=1)
        grp, gci, gva = f.values()
        G = sp.csr_matrix((of.bf16_round(gva.astype(np.float32)).astype(np.float64), gci, grp),
                          shape=(m, m))
        z0 = G @ (G.T @ b.astype(np.float32).astype(np.float64))
        err = np.linalg.norm(x / np.linalg.norm(x) - z0 / np.linalg.norm(z0))
        assert err <= 1e-5, err
        assert st.preconditioner == 2 and st.iterations == 1
    finally:
        f.release()
        net.release()


@pytest.mark.parametrize("x64,kind", [("0", "3d"), ("1", "3d"), ("0", "2d")])
def test_fused_update_iteration_matches_the_separate_kernels(ctl, monkeypatch, x64, kind):
    """With the factor in diagonal storage the x / r update is fused into
    w = G^T r (r alternating between two buffers): the iterates, iteration
    and replacement counts equal those of the separate kernels bit for bit,
    through residual replacements after chunks cut short by convergence
    (after an odd count the residual is moved back to its buffer before the
    replacement reads it), with the solution as an FP32 increment and in
    FP64."""
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_POLICY", "always")
    monkeypatch.setenv("CLSPARSE_AMD_MIXED_X64", x64)
    from clsparse_amd import fsai as fs
    params, _ = fs.load_weights()
    if kind == "3d":
        rp, ci, v, m = cls.gen_diffusion3d(28, 26, 24, sigma=1.0, ax=1.0, ay=1.0, az=100.0, seed=6)
    else:
        rp, ci, v, m = _aniso(130, 120, seed=6)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    net = cls.Mlp(seed=1)
    net.set_params(params)
    f = cls.Fsai(A, ctl)
    try:
        f.apply_net(net, ctl)
        assert len(f.diagonals()) == (4 if kind == "3d" else 5)
        b = cls.gen_uniform(m, -1, 1, seed=31)
        out = {}
        for fuse in ("1", "0"):
            monkeypatch.setenv("CLSPARSE_AMD_CG_FUSE", fuse)
            say(f"fsai: fused update {fuse}, x64 {x64}, {kind}")
            out[fuse] = [_solve_fsai(ctl, A, f, b, tol=tol) for tol in (1e-10, 3e-9, 1e-9, 3e-10)]
        for (s1, x1), (s0, x0) in zip(out["1"], out["0"]):
            assert s1.preconditioner == 2 and s0.preconditioner == 2
            assert (s1.iterations, s1.outer_iterations) == (s0.iterations, s0.outer_iterations)
            assert s1.outer_iterations >= 2
            np.testing.assert_array_equal(x1, x0)
    finally:
        f.release()
        net.release()


def test_first_iteration_applies_an_irregular_factor(ctl, monkeypatch):
    """An irregular pattern (columns of G from 1 to dozens of entries) keeps
    the CSR SpMVs of G and G^T: one PCG iteration from x = 0 is x = alpha
    G (G^T b) with the BF16-rounded factor, so G^T's row pointer, columns and
    value positions (from the device sort of G's entries) are right."""
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_POLICY", "always")
    from clsparse_amd import fsai as fs
    params, _ = fs.load_weights()
    rng = np.random.default_rng(11)
    a = sp.random(3000, 3000, density=0.004, random_state=rng, format="csr")
    a = a + a.T
    a = a + sp.diags(np.asarray(abs(a).sum(axis=1)).ravel() + 1.0)
    a = sp.csr_matrix(a)
    a.sort_indices()
    rp, ci, v, m = a.indptr.astype(np.uint32), a.indices.astype(np.uint32), a.data, 3000
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    net = cls.Mlp(seed=1)
    net.set_params(params)
    f = cls.Fsai(A, ctl)
    try:
        f.apply_net(net, ctl)
        assert f.diagonals() == []
        grp, gci, gva = f.values()
        counts = np.bincount(gci.astype(np.int64), minlength=m)
        assert counts.min() >= 1 and counts.max() >= 10
        b = cls.gen_uniform(m, -1, 1, seed=13)
        say("fsai: first iteration, irregular CSR factor")
        st, x = _solve_fsai(ctl, A, f, b, max_iters=1)
        G = sp.csr_matrix((of.bf16_round(gva.astype(np.float32)).astype(np.float64), gci, grp),
                          shape=(m, m))
        z0 = G @ (G.T @ b.astype(np.float32).astype(np.float64))
        err = np.linalg.norm(x / np.linalg.norm(x) - z0 / np.linalg.norm(z0))
        assert err <= 1e-5, err
        assert st.preconditioner == 2 and st.iterations == 1
    finally:
        f.release()
        net.release()


@pytest.mark.parametrize("dims", [(24, 24, 24), (64, 64, 16)])
def test_diagonal_storage_solves_like_the_csr_factor(ctl, monkeypatch, dims):
    """The learned factor on a 3-D anisotropic stencil: the solve from
    diagonal storage and the one through the CSR SpMVs of G and G^T take the
    same iterations (the row sums are formed in the same order; only a
    matrix's last row block may sum in another order) and both reach 1e-8."""
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_POLICY", "always")
    from clsparse_amd import fsai as fs
    params, _ = fs.load_weights()
    # 64 x 64 x 16: A's largest offset is 4096 rows, so the FP32 SpMV of A
    # walks rows in line order with the far operands carried in registers.
    rp, ci, v, m = cls.gen_diffusion3d(*dims, sigma=1.0, ax=1.0, ay=1.0, az=100.0, seed=5)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    net = cls.Mlp(seed=1)
    net.set_params(params)
    f = cls.Fsai(A, ctl)
    Acpu = sp.csr_matrix((v, ci, rp), shape=(m, m))
    try:
        f.apply_net(net, ctl)
        b = cls.gen_uniform(m, -1, 1, seed=13)
        say(f"fsai: diagonal storage vs CSR factor, {dims}")
        monkeypatch.setenv("CLSPARSE_AMD_FSAI_DIA", "0")
        st_c, x_c = _solve_fsai(ctl, A, f, b)
        monkeypatch.setenv("CLSPARSE_AMD_FSAI_DIA", "1")
        st_d, x_d = _solve_fsai(ctl, A, f, b)
        assert st_c.status_detail == 0 and st_d.status_detail == 0
        assert abs(st_c.iterations - st_d.iterations) <= max(2, st_c.iterations // 50)
        for x in (x_c, x_d):
            assert np.linalg.norm(b - Acpu @ x) / np.linalg.norm(b) <= 1e-8
    finally:
        f.release()
        net.release()


@pytest.mark.parametrize("dia", ["1", "0"])
def test_factor_from_unsorted_rows_applies_like_the_sorted_one(ctl, monkeypatch, dia):
    """The pattern of G and G^T is built on the device from A's rows in the
    order they are stored: with the columns of every row shuffled, G keeps
    A's order within a row, holds the same entries, and one PCG iteration
    applies the same G G^T (diagonal storage and the CSR SpMVs, whose G^T
    is the device-sorted transpose)."""
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_POLICY", "always")
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_DIA", dia)
    from clsparse_amd import fsai as fs
    params, _ = fs.load_weights()
    rp, ci, v, m = cls.gen_diffusion3d(13, 11, 9, sigma=1.0, ax=1.0, ay=1.0, az=100.0, seed=8)
    rng = np.random.default_rng(8)
    perm = np.concatenate([rp[r] + rng.permutation(rp[r + 1] - rp[r]) for r in range(m)])
    ci_s, v_s = ci[perm], v[perm]
    A = cls.CsrMatrix(rp, ci_s, v_s, m)
    A.meta_create(ctl)
    net = cls.Mlp(seed=1)
    net.set_params(params)
    f = cls.Fsai(A, ctl)
    try:
        f.apply_net(net, ctl)
        assert len(f.diagonals()) == 4
        grp, gci, gva = f.values()
        rows = np.repeat(np.arange(m, dtype=np.int64), np.diff(rp.astype(np.int64)))
        lower = ci_s.astype(np.int64) <= rows
        np.testing.assert_array_equal(gci, ci_s[lower])  # A's order within each row
        np.testing.assert_array_equal(np.diff(grp.astype(np.int64)),
                                      np.bincount(rows[lower], minlength=m))
        Gs = sp.csr_matrix((gva, gci, grp), shape=(m, m))
        Ao = cls.CsrMatrix(rp, ci, v, m)
        fo = cls.Fsai(Ao, ctl)
        try:
            fo.apply_net(net, ctl)
            Go = sp.csr_matrix(tuple(reversed(fo.values())), shape=(m, m))
            assert abs(Gs - Go).max() == 0.0
        finally:
            fo.release()
        b = cls.gen_uniform(m, -1, 1, seed=21)
        say(f"fsai: unsorted rows, first iteration, diagonal storage={dia}")
        st, x = _solve_fsai(ctl, A, f, b, max_iters=1)
        G = sp.csr_matrix((of.bf16_round(gva.astype(np.float32)).astype(np.float64), gci, grp),
                          shape=(m, m))
        z0 = G @ (G.T @ b.astype(np.float32).astype(np.float64))
        err = np.linalg.norm(x / np.linalg.norm(x) - z0 / np.linalg.norm(z0))
        assert err <= 1e-5, err
        assert st.preconditioner == 2 and st.iterations == 1
    finally:
        f.release()
        net.release()


def test_factor_rejects_bad_row_pointer_and_columns(ctl):
    """clsparseAmdFsaiCreate checks A's structure on the device: a row
    pointer that decreases or passes nnz, or a column index >= n, is
    clsparseInvalidValue (the kernels read only inside the arrays)."""
    rp, ci, v, m = cls.gen_diffusion3d(6, 5, 4, sigma=0.0, ax=1.0, ay=1.0, az=1.0, seed=1)
    bad_rp = rp.copy()
    bad_rp[7] = bad_rp[8] + 1  # row 6 ends before it starts
    bad_end = rp.copy()
    bad_end[-1] = rp[-1] + 3  # last row runs past nnz
    bad_end[-2] = min(bad_end[-2], rp[-1])
    bad_ci = ci.copy()
    bad_ci[10] = m + 5
    for k, (r, c) in enumerate([(bad_rp, ci), (bad_end, ci), (rp, bad_ci)]):
        A = cls.CsrMatrix(r, c, v, m)
        say(f"fsai: bad structure case {k}")
        with pytest.raises(cls.ClsparseError):
            cls.Fsai(A, ctl)
    A = cls.CsrMatrix(rp, ci, v, m)
    f = cls.Fsai(A, ctl)  # the same matrix with a valid structure
    f.release()


def test_factor_rejects_sizes_beyond_int_counts(ctl):
    """The device scans and sorts of the pattern count in int: n or nnz of
    2^31 or more is clsparseInvalidSize, before any array is touched."""
    A = cls.clsparseCsrMatrix()
    cls.load().clsparseInitCsrMatrix(ctypes.byref(A))
    for rows, nnz in ((1 << 31, 10), (1000, 1 << 31)):
        A.num_rows = A.num_cols = rows
        A.num_nonzeros = nnz
        h = ctypes.c_void_p()
        say(f"fsai: size guard rows={rows} nnz={nnz}")
        assert cls.load().clsparseAmdFsaiCreate(ctypes.byref(h), ctypes.byref(A), ctl.handle) == \
            cls.clsparseInvalidSize
