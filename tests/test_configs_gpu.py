"""The solver and training workloads at their full sizes, judged on the CPU.

* 128^3 variable-coefficient diffusion (n = 2,097,152, nnz = 14,581,760): the
  mixed-precision solver (FP32 inner CG, FP64 residual replacement) reaches an
  FP64 true relative residual <= 1e-10, with A's SpMV from diagonal storage
  and from CSR-Adaptive.
* 160^3 held-out anisotropic diffusion (n = 4,096,000, nnz = 28,518,400):
  learned-FSAI mixed PCG with the committed network, and FP64 Jacobi-PCG, each
  to an FP64 true relative residual <= 1e-8; the learned solve takes fewer
  iterations; the learned solve with A, G and G^T on CSR-Adaptive as well.
* 512^2 training family (n = 262,144): one training step's Hutchinson loss and
  dL/dG against oracle/fsai.py.
* The Adam update against a step computed by hand from the same gradient.

Every residual is b - A x formed by the oracle's compensated CSR product on the
host from the solution the library returned (never by the library itself);
the cited reference solver test runs the drop-in on a full user matrix the
same way (src/tests/test-solvers.cpp L123-237).
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import clsparse_amd as cls
import oracle
from clsparse_amd import fsai as fs
from oracle import fsai as of

from gpu_helpers import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctl():
    c = cls.Control()
    yield c
    c.release()


def _true_rel_residual(rp, ci, v, x, b):
    ax, _ = oracle.csrmv(rp, ci, v, np.asarray(x, dtype=np.float64))
    return float(np.linalg.norm(b - ax) / np.linalg.norm(b))


def _solve(ctl, fn, A, b, tol, f=None, max_iters=50000):
    n = A.num_rows
    xd = cls.DenseVector(torch.zeros(n, dtype=torch.float64, device="cuda"))
    bd = cls.DenseVector(torch.tensor(b, dtype=torch.float64, device="cuda"))
    sc = cls.SolverControl(cls.DIAGONAL, max_iters, tol, 0.0)
    try:
        args = (ctypes.byref(xd.s), ctypes.byref(A.s), ctypes.byref(bd.s))
        args += ((f.handle,) if f is not None else ()) + (sc.handle, ctl.handle)
        cls.check("solve", fn(*args))
        return sc.stats(), xd.t.cpu().numpy()
    finally:
        sc.release()


@pytest.fixture(scope="module")
def diffusion128():
    rp, ci, v, m = cls.gen_diffusion3d(128, 128, 128, sigma=1.0, seed=3)
    b = cls.gen_uniform(m, -1.0, 1.0, seed=4)
    assert m == 2097152 and int(rp[-1]) == 14581760
    return rp, ci, v, m, b


@pytest.mark.parametrize("x64", ["0", "1"], ids=["fp32-increment", "fp64-x"])
@pytest.mark.parametrize("dia", ["1", "0"], ids=["diagonal-storage", "csr-adaptive"])
def test_mixed_solver_128cubed_to_1e10(ctl, monkeypatch, diffusion128, dia, x64):
    """x64 = 1: the solution is updated in FP64 every iteration
    (CLSPARSE_AMD_MIXED_X64) instead of through an FP32 increment."""
    monkeypatch.setenv("CLSPARSE_AMD_DIA", dia)
    monkeypatch.setenv("CLSPARSE_AMD_MIXED_X64", x64)
    rp, ci, v, m, b = diffusion128
    say(f"128^3 mixed solver to 1e-10, dia={dia} x64={x64}")
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    st, x = _solve(ctl, cls.load().clsparseAmdDcsrcgMixed, A, b, 1e-10)
    A.meta_delete()
    assert st.status_detail == 0
    assert st.a_diagonals == (7 if dia == "1" else 0)
    res = _true_rel_residual(rp, ci, v, x, b)
    say(f"  iterations {st.iterations}, oracle residual {res:.3e}, library {st.true_residual:.3e}")
    assert res <= 1e-10 * (1 + 1e-6)
    assert abs(res - st.true_residual) <= 1e-3 * res


@pytest.fixture(scope="module")
def target160():
    rp, ci, v, m = fs.gen_target3d(160)
    b = cls.gen_uniform(m, -1.0, 1.0, seed=41)
    assert m == 4096000 and int(rp[-1]) == 28518400
    return rp, ci, v, m, b


@pytest.fixture(scope="module")
def net():
    params, _ = fs.load_weights()
    n = cls.Mlp(seed=1)
    n.set_params(params)
    yield n
    n.release()


def test_learned_fsai_and_fp64_jacobi_160cubed_to_1e8(ctl, monkeypatch, target160, net):
    monkeypatch.delenv("CLSPARSE_AMD_FSAI_POLICY", raising=False)
    rp, ci, v, m, b = target160
    lib = cls.load()
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    f = cls.Fsai(A, ctl)
    try:
        f.apply_net(net, ctl)
        say("160^3 learned-FSAI mixed PCG to 1e-8")
        st_l, x_l = _solve(ctl, lib.clsparseAmdDcsrcgFsai, A, b, 1e-8, f=f)
        res_l = _true_rel_residual(rp, ci, v, x_l, b)
        say(f"  iterations {st_l.iterations}, preconditioner {st_l.preconditioner}, "
            f"oracle residual {res_l:.3e}")
        say("160^3 FP64 Jacobi-PCG to 1e-8")
        st_j, x_j = _solve(ctl, lib.clsparseAmdDcsrcgRefined, A, b, 1e-8)
        res_j = _true_rel_residual(rp, ci, v, x_j, b)
        say(f"  iterations {st_j.iterations}, oracle residual {res_j:.3e}")
    finally:
        f.release()
        A.meta_delete()
    assert st_l.status_detail == 0 and st_j.status_detail == 0
    assert st_l.preconditioner == 2  # the learned factor, not the Jacobi fallback
    assert res_l <= 1e-8 * (1 + 1e-6) and res_j <= 1e-8 * (1 + 1e-6)
    # What the network gains: fewer iterations than Jacobi on the held-out problem.
    assert st_l.iterations < 0.75 * st_j.iterations, (st_l.iterations, st_j.iterations)


@pytest.mark.parametrize("bf16,x64", [("1", "0"), ("0", "1")], ids=["bf16-g", "fp32-g-fp64-x"])
def test_learned_fsai_160cubed_on_csr_adaptive(ctl, monkeypatch, target160, net, bf16, x64):
    """A, G and G^T through CSR-Adaptive SpMVs (no diagonal storage); G's
    values in BF16 or FP32, the solution as an FP32 increment or in FP64."""
    monkeypatch.setenv("CLSPARSE_AMD_DIA", "0")
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_DIA", "0")
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_POLICY", "always")
    monkeypatch.setenv("CLSPARSE_AMD_FSAI_BF16", bf16)
    monkeypatch.setenv("CLSPARSE_AMD_MIXED_X64", x64)
    rp, ci, v, m, b = target160
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    f = cls.Fsai(A, ctl)
    try:
        f.apply_net(net, ctl)
        say("160^3 learned-FSAI mixed PCG on CSR-Adaptive to 1e-8")
        st, x = _solve(ctl, cls.load().clsparseAmdDcsrcgFsai, A, b, 1e-8, f=f)
    finally:
        f.release()
        A.meta_delete()
    res = _true_rel_residual(rp, ci, v, x, b)
    say(f"  iterations {st.iterations}, oracle residual {res:.3e}")
    assert st.status_detail == 0 and st.preconditioner == 2
    assert st.a_diagonals == 0 and st.g_diagonals == 0
    assert res <= 1e-8 * (1 + 1e-6)


def test_training_step_512_family_matches_oracle(ctl, net):
    """One step on a matrix of the 512^2 training family: Hutchinson loss and
    dL/dG per nonzero of G against the oracle, from the device's network output."""
    rp, ci, v, m = fs.gen_family(3)
    assert m == 512 * 512
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    f = cls.Fsai(A, ctl)
    probes, seed = 8, 1234
    try:
        say("512^2 training step vs oracle")
        loss, gG = f.train_step(net, A, ctl, probes=probes, seed=seed, lr=0.0, want_grad=True)
        o = f.output().astype(np.float64)
    finally:
        f.release()
        A.meta_delete()
    Acpu = sp.csr_matrix((v, ci, rp), shape=(m, m))
    G = of.assemble_g(rp, ci, v, o)
    Z = np.stack([of.rademacher(m, of.probe_seed(seed, k)) for k in range(probes)], axis=1)
    loss_ref, grad_ref = of.hutchinson(Acpu, G, Z)
    say(f"  loss {loss:.6e} oracle {loss_ref:.6e}")
    assert loss == pytest.approx(loss_ref, rel=1e-11)
    gi, gj, _ = of.tril_pattern(rp, ci)
    gref = np.asarray(grad_ref[gi, gj]).ravel()
    np.testing.assert_allclose(gG, gref, rtol=1e-9, atol=1e-12 * np.abs(gref).max())


def _adam_ref(p, m1, m2, g, lr, step):
    """One Adam step (beta 0.9 / 0.999, eps 1e-8) in FP32, as written out in
    Kingma and Ba, Algorithm 1, with the bias corrections of step `step`."""
    f = np.float32
    g = g.astype(f)
    m1 = f(0.9) * m1 + f(0.1) * g
    m2 = f(0.999) * m2 + f(0.001) * g * g
    mh = m1 / (f(1) - f(0.9) ** f(step))
    vh = m2 / (f(1) - f(0.999) ** f(step))
    return p - f(lr) * mh / (np.sqrt(vh) + f(1e-8)), m1, m2


def test_adam_steps_match_hand_computed_update(ctl):
    """Two training steps with lr > 0: the parameters after each equal the
    Adam update of the step's own gradient (read back from the library)
    computed here in FP32, within a few FP32 roundings of the step."""
    rp, ci, v, m = cls.gen_aniso2d(24, 20, corr_len=4.0, sigma=1.0, ratio=100.0, theta=0.6, seed=3)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    rng = np.random.default_rng(5)
    p0 = np.zeros(of.N_PARAMS, dtype=np.float32)
    p0[:] = rng.normal(scale=0.1, size=of.N_PARAMS)
    nt = cls.Mlp(seed=9)
    f = cls.Fsai(A, ctl)
    lr = 1e-3
    try:
        nt.set_params(p0)
        m1 = np.zeros(of.N_PARAMS, np.float32)
        m2 = np.zeros(of.N_PARAMS, np.float32)
        p = p0.copy()
        for step in (1, 2):
            say(f"adam step {step}")
            f.train_step(nt, A, ctl, probes=4, seed=40 + step, lr=lr)
            g = nt.grad()
            p_ref, m1, m2 = _adam_ref(p, m1, m2, g, lr, step)
            p_dev = nt.params()
            # each update moves a parameter by at most ~lr; allow FP32 rounding
            # of the update and of the parameter (contractions may differ)
            tol = 1e-5 * lr + 2.0 ** -23 * np.abs(p_ref)
            bad = np.abs(p_dev - p_ref) > tol
            assert not bad.any(), (int(np.argmax(bad)), float(p_dev[bad][0]), float(p_ref[bad][0]))
            assert np.abs(p_dev - p).max() > 0.5 * lr  # the step did move the parameters
            p = p_dev.astype(np.float32)
    finally:
        f.release()
        nt.release()
        A.meta_delete()
