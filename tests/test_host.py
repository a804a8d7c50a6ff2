"""CPU-side checks of the product library: ABI, host binning, generators.

The HIP library loads without a GPU; only entry points that never touch the
device are called here.
"""
import ctypes
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import clsparse_amd as cls
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "row_blocks_known.json")


def _declared_symbols():
    names = set()
    for h in ("clSPARSE.h", "clSPARSE-amd-ext.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        for m in re.finditer(r"CLSPARSE_EXPORT\s+[\w\s\*]+?\b(\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = cls.load()
    declared = _declared_symbols()
    assert len(declared) >= 57
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(cls.EXPORTED)


def test_header_struct_layout_matches_python_mirror(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "clSPARSE.h"
#include "clSPARSE-amd-ext.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(clsparseScalar), sizeof(cldenseVector),
         sizeof(clsparseCsrMatrix), offsetof(clsparseCsrMatrix, meta),
         sizeof(clsparseAmdSolverStats), sizeof(clsparseAmdHostCsr));
  return 0;
}''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    got = [int(v) for v in out]
    want = [ctypes.sizeof(cls.clsparseScalar), ctypes.sizeof(cls.cldenseVector),
            ctypes.sizeof(cls.clsparseCsrMatrix), cls.clsparseCsrMatrix.meta.offset,
            ctypes.sizeof(cls.clsparseAmdSolverStats), ctypes.sizeof(cls.clsparseAmdHostCsr)]
    assert got == want
    # LP64 layout of the reference structs with cl_mem -> pointer.
    assert got[:4] == [16, 24, 64, 56]


def test_version_and_setup_without_device():
    lib = cls.load()
    v = [ctypes.c_uint32() for _ in range(4)]
    assert lib.clsparseGetVersion(*[ctypes.byref(t) for t in v]) == 0
    assert [t.value for t in v] == [0, 10, 2, 0]
    assert lib.clsparseTeardown() == 0
    # Entry points check the init flag before touching the device.
    s = cls.clsparseScalar()
    x = cls.cldenseVector()
    assert lib.cldenseDnrm1(ctypes.byref(s), ctypes.byref(x), ctypes.c_void_p(1)) == \
        cls.clsparseNotInitialized
    assert lib.clsparseSetup() == 0
    assert lib.cldenseDnrm1(ctypes.byref(s), ctypes.byref(x), None) == \
        cls.clsparseInvalidControlObject
    assert lib.clsparseReleaseSolverControl(None) == cls.clsparseInvalidSolverControlObject
    r = lib.clsparseCreateSolverControl(cls.DIAGONAL, 100, 1e-2, 1e-5)
    assert r.status == 0
    st = cls.clsparseAmdSolverStats()
    assert lib.clsparseAmdGetSolverStats(ctypes.c_void_p(r.control), ctypes.byref(st)) == 0
    assert (st.max_iters, st.relative_tolerance, st.absolute_tolerance) == (100, 1e-2, 1e-5)
    assert lib.clsparseReleaseSolverControl(ctypes.c_void_p(r.control)) == 0


@pytest.mark.parametrize("case", json.load(open(GOLDEN))["cases"], ids=lambda c: c["name"])
def test_library_binning_known_answers(case):
    rp = np.concatenate([[0], np.cumsum(case["row_lengths"])]).astype(np.uint32)
    entries, size = cls.row_blocks_host(rp)
    assert [int(e) for e in entries] == case["entries"]
    assert size == case["row_block_size"]
    assert cls.row_blocks_host(rp, count_only=True) == case["size_pass"]


@pytest.mark.parametrize("seed", range(8))
def test_library_binning_matches_oracle_random(seed):
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(1, 50000))
    mix = [rng.integers(0, 8, n), rng.integers(0, 300, n),
           np.clip(np.round(np.exp(rng.normal(1.6, 1.5, n))), 0, 200000),
           rng.choice([0, 1, 2, 31, 32, 33, 127, 128, 129, 1023, 1024, 1025, 3072, 3073, 9000], n)]
    lens = mix[seed % 4].astype(np.int64)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    e1, s1 = cls.row_blocks_host(rp)
    e2, s2 = oracle.row_blocks(rp)
    assert s1 == s2
    np.testing.assert_array_equal(e1, e2)
    assert cls.row_blocks_host(rp, count_only=True) == oracle.row_blocks(rp, count_only=True)


def test_library_binning_skips_match_oracle():
    """The binning skips runs of rows of at most 128 nonzeros outside long-row
    regions; patterns that stress where the skip stops: the full power-law
    workload, regular stencil rows, long runs of empty rows, long-row regions
    continued by 32..128-nonzero rows and left by shorter ones."""
    rng = np.random.default_rng(77)
    pats = [cls.gen_powerlaw(1 << 21, 16.0, 1.5, 200000, 8, 100000, seed=1)[0],
            cls.gen_diffusion3d(96, 96, 96)[0],
            np.concatenate([[0], np.cumsum(np.tile([200, 40, 129, 31, 0, 0, 5, 128, 1024, 3073, 32, 33], 5000))]),
            np.concatenate([[0], np.cumsum(np.where(rng.random(300000) < 0.01, rng.integers(0, 3000, 300000),
                                                    0))]),
            np.concatenate([[0], np.cumsum(rng.choice([0, 1, 127, 128, 129, 1000, 1023, 1024, 1025], 400000,
                                                      p=[.3, .3, .1, .1, .05, .05, .04, .03, .03]))]),
            # Large enough to be binned in several ranges on several threads,
            # cut at long rows after short ones (lengths 129, 1024, 2000, 3073).
            np.concatenate([[0], np.cumsum(np.tile([200, 40, 129, 31, 0, 0, 5, 128, 1024, 3073, 32, 33], 60000))]),
            np.concatenate([[0], np.cumsum(np.tile([5, 1024, 3, 2000, 10, 129, 0, 1024, 7, 7, 7], 90000))]),
            np.concatenate([[0], np.cumsum(rng.choice([0, 3, 9, 31, 129, 300, 1024, 5000], 1500000,
                                                      p=[.2, .3, .3, .1, .04, .03, .02, .01]))])]
    for rp in pats:
        rp = np.ascontiguousarray(rp, dtype=np.uint32)
        e1, s1 = cls.row_blocks_host(rp)
        e2, s2 = oracle.row_blocks(rp)
        assert s1 == s2
        np.testing.assert_array_equal(e1, e2)
        assert cls.row_blocks_host(rp, count_only=True) == oracle.row_blocks(rp, count_only=True)


def test_library_binning_regime_changes_match_oracle():
    # Runs of equal-length rows whose length changes abruptly (empty rows,
    # stencil-like rows, rows near the 128 / 1024 limits, long rows): a block
    # after a change holds a different number of rows than the one before,
    # which the skip's search starts from.
    rng = np.random.default_rng(17)
    choices = np.array([0, 1, 3, 7, 31, 32, 64, 127, 128, 129, 300, 1023, 1024, 1025, 5000])
    for trial in range(6):
        runs = []
        while sum(len(r) for r in runs) < 400000:
            runs.append(np.full(int(rng.integers(1, 3000)), int(rng.choice(choices))))
            if rng.random() < 0.3:
                runs.append(rng.integers(0, 20, int(rng.integers(1, 500))))
        lens = np.concatenate(runs)[:400000]
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
        e1, s1 = cls.row_blocks_host(rp)
        e2, s2 = oracle.row_blocks(rp)
        assert s1 == s2, trial
        np.testing.assert_array_equal(e1, e2)
        assert cls.row_blocks_host(rp, count_only=True) == oracle.row_blocks(rp, count_only=True)


def test_library_binning_counts_a_decreasing_row_pointer_as_the_reference():
    # clsparseCsrMetaSize counts any row pointer as the reference's pass does;
    # a decreasing one takes the sequential pass (the search needs it sorted).
    rng = np.random.default_rng(5)
    small = np.array([0, 5, 10, 8, 20, 20, 400, 399, 2000, 2001], dtype=np.uint32)
    big = np.concatenate([[0], np.cumsum(rng.integers(0, 40, 600000))]).astype(np.uint32)
    big[300001] = big[300000] - 3
    for rp in (small, big):
        assert cls.row_blocks_host(rp, count_only=True) == oracle.row_blocks(rp, count_only=True)


_FORK_SCRIPT = r"""
import os, sys
import numpy as np
import clsparse_amd as cls
rng = np.random.default_rng(9)
rp = np.concatenate([[0], np.cumsum(rng.integers(0, 300, 1 << 20))]).astype(np.uint32)
ref, size = cls.row_blocks_host(rp)  # the parent's threads start
pid = os.fork()
if pid == 0:
    e, s = cls.row_blocks_host(rp)
    sys.exit(0 if (s == size and np.array_equal(e, ref)) else 3)  # a normal exit: destructors run
_, status = os.waitpid(pid, 0)
assert os.WIFEXITED(status) and os.WEXITSTATUS(status) == 0, status
e, s = cls.row_blocks_host(rp)
assert s == size and np.array_equal(e, ref)
print("ok")
"""


def test_library_binning_in_a_forked_child():
    # The binning's host threads start on first use; a child of fork()
    # inherits none of them and must bin on its own thread (and exit) rather
    # than wait for them. Run in a process of its own, with a time limit.
    r = subprocess.run([sys.executable, "-c", _FORK_SCRIPT], cwd=ROOT, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout[-2000:], r.stderr[-2000:])


def test_library_binning_config_patterns_match_oracle():
    # The three structured workloads at full size (pattern only).
    for rp in (cls.gen_poisson2d(256, 256)[0], cls.gen_diffusion3d(64, 64, 64)[0]):
        e1, s1 = cls.row_blocks_host(rp)
        e2, s2 = oracle.row_blocks(rp)
        assert s1 == s2
        np.testing.assert_array_equal(e1, e2)


def test_poisson_generator():
    rp, ci, v, m = cls.gen_poisson2d(256, 256)
    n = 256 * 256
    assert m == n and rp[-1] == 5 * n - 4 * 256 == 326656
    # sorted unique columns, diagonal 4, symmetric
    for r in (0, 1, 255, 256, 257, 30000, n - 1):
        cols = ci[rp[r]:rp[r + 1]]
        assert np.all(np.diff(cols.astype(np.int64)) > 0)
        assert v[rp[r] + list(cols).index(r)] == 4.0
    import scipy.sparse as sp
    a = sp.csr_matrix((v, ci, rp), shape=(n, n))
    assert abs(a - a.T).max() == 0


def test_diffusion_generator_spd_structure():
    import scipy.sparse as sp
    rp, ci, v, m = cls.gen_diffusion3d(12, 10, 8, sigma=1.0, ax=1.0, ay=1.0, az=0.01, seed=5)
    n = 12 * 10 * 8
    assert rp[-1] == 7 * n - 2 * (12 * 10 + 10 * 8 + 12 * 8)
    a = sp.csr_matrix((v, ci, rp), shape=(n, n))
    assert abs(a - a.T).max() == 0  # exactly symmetric (the face coefficient is)
    d = a.diagonal()
    off = np.asarray(abs(a).sum(axis=1)).ravel() - d
    # Weakly diagonally dominant; strictly on rows with a Dirichlet face.
    assert np.all(d >= off * (1 - 1e-14))
    boundary = np.diff(rp.astype(np.int64)) < 7
    assert np.all(d[boundary] > off[boundary])
    assert np.linalg.eigvalsh(a.toarray()).min() > 0


def _aniso_stencil(eps):
    """Closed-form Q1 stencil of -div(diag(1, eps) grad u), unit cells:
    the x part is (1-D stiffness [-1 2 -1]) x (1-D mass [1/6 2/3 1/6]) and the
    y part its transpose scaled by eps."""
    s1, m1 = np.array([-1.0, 2.0, -1.0]), np.array([1 / 6, 2 / 3, 1 / 6])
    return np.outer(m1, s1) + eps * np.outer(s1, m1)  # [dy][dx]


@pytest.mark.parametrize("theta,ratio", [(0.0, 1.0), (0.0, 50.0), (np.pi / 2, 50.0)])
def test_aniso2d_generator_closed_form_stencil(theta, ratio):
    import scipy.sparse as sp
    nx, ny = 7, 6
    rp, ci, v, m = cls.gen_aniso2d(nx, ny, corr_len=4.0, sigma=0.0, ratio=ratio, theta=theta,
                                   seed=2)
    n = nx * ny
    assert m == n and rp[-1] == 9 * n - 6 * (nx + ny) + 4
    a = sp.csr_matrix((v, ci, rp), shape=(n, n)).toarray()
    st = _aniso_stencil(1.0 / ratio)
    if theta != 0.0:
        st = st.T  # rotating by 90 degrees swaps the strong and the weak axis
    # every interior node: the 3x3 stencil, exactly (up to rounding of cos/sin)
    for y in range(ny):
        for x in range(nx):
            i = y * nx + x
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    xx, yy = x + dx, y + dy
                    if 0 <= xx < nx and 0 <= yy < ny:
                        assert abs(a[i, yy * nx + xx] - st[dy + 1, dx + 1]) < 1e-12
    if ratio == 1.0:  # isotropic Q1 Laplacian: 8/3 and -1/3
        assert abs(st[1, 1] - 8 / 3) < 1e-15 and abs(st[0, 0] + 1 / 3) < 1e-15
        assert abs(st[0, 1] + 1 / 3) < 1e-15


def test_aniso2d_generator_spd_random_field():
    import scipy.sparse as sp
    nx, ny = 16, 12
    rp, ci, v, m = cls.gen_aniso2d(nx, ny, corr_len=4.0, sigma=1.0, ratio=300.0, theta=0.6,
                                   seed=9)
    n = nx * ny
    a = sp.csr_matrix((v, ci, rp), shape=(n, n))
    assert abs(a - a.T).max() <= 1e-14 * abs(a).max()
    assert np.linalg.eigvalsh(a.toarray()).min() > 0
    for r in range(n):
        c = ci[rp[r]:rp[r + 1]].astype(np.int64)
        assert np.all(np.diff(c) > 0)
    # the coefficient field varies (sigma > 0) and the generator is seeded
    d = a.diagonal()
    assert d.max() / d.min() > 2.0
    rp2, ci2, v2, _ = cls.gen_aniso2d(nx, ny, corr_len=4.0, sigma=1.0, ratio=300.0, theta=0.6,
                                      seed=9)
    np.testing.assert_array_equal(v, v2)
    _, _, v3, _ = cls.gen_aniso2d(nx, ny, corr_len=4.0, sigma=1.0, ratio=300.0, theta=0.6, seed=10)
    assert not np.array_equal(v, v3)


def test_powerlaw_generator_statistics():
    rp, ci, v, m = cls.gen_powerlaw(1 << 16, 16.0, 1.5, 200000, 2, 20000, seed=3)
    lens = np.diff(rp.astype(np.int64))
    assert lens.min() >= 1 and (lens == 20000).sum() >= 2
    body = lens[lens != 20000]
    assert 14.0 < body.mean() < 18.0
    for r in range(0, m, 997):
        c = ci[rp[r]:rp[r + 1]].astype(np.int64)
        assert np.all(np.diff(c) > 0) and c.min() >= 0 and c.max() < m
    assert np.all(np.abs(v) <= 1.0)
    # deterministic in the seed
    rp2, ci2, v2, _ = cls.gen_powerlaw(1 << 16, 16.0, 1.5, 200000, 2, 20000, seed=3)
    np.testing.assert_array_equal(ci, ci2)
    np.testing.assert_array_equal(v, v2)


def test_uniform_generator():
    a = cls.gen_uniform(100000, -1.0, 1.0, seed=11)
    assert a.min() >= -1.0 and a.max() < 1.0 and abs(a.mean()) < 0.01
    np.testing.assert_array_equal(a, cls.gen_uniform(100000, -1.0, 1.0, seed=11))
    assert not np.array_equal(a, cls.gen_uniform(100000, -1.0, 1.0, seed=12))


def test_fsai_weights_roundtrip_and_checksum(tmp_path):
    from clsparse_amd import fsai as fs
    n = cls.load().clsparseAmdMlpNumParams()
    assert n == 64 * 16 + 64 + 64 * 64 + 64 + 64 + 1
    p = np.random.default_rng(0).normal(size=n).astype(np.float32)
    path = str(tmp_path / "w.npy")
    digest = fs.save_weights(path, p, {"note": "test"})
    q, rec = fs.load_weights(path)
    np.testing.assert_array_equal(p, q)
    assert rec["sha256_float32_le"] == digest and rec["note"] == "test"
    p[7] += 1.0
    np.save(path, p, allow_pickle=False)
    with pytest.raises(ValueError):
        fs.load_weights(path)


def test_fsai_training_family_is_seeded():
    from clsparse_amd import fsai as fs
    a, b = fs.aniso2d_params(5), fs.aniso2d_params(5)
    assert a == b and a != fs.aniso2d_params(6)
    ratios = [fs.aniso2d_params(i)["ratio"] for i in range(400)]
    assert 1.0 <= min(ratios) and max(ratios) <= 1000.0
    assert 0.3 < np.mean(np.log10(ratios)) / 1.5 < 1.7


def test_committed_fsai_weights_load_and_checksum():
    from clsparse_amd import fsai as fs
    p, rec = fs.load_weights()
    assert p.shape == (cls.load().clsparseAmdMlpNumParams(),) and np.all(np.isfinite(p))
    assert rec["held_out_loss_learned_mean"] < rec["held_out_loss_jacobi_mean"]
    assert rec["train"]["matrices"] == 1800


def test_line_order_partition_covers_every_row_once(tmp_path):
    """The line-order row partition of the diagonal-storage kernels
    (clsparse_amd/csrc/line_part.hpp), walked on the CPU exactly as the
    kernels walk it: every row of the matrix is visited once, for grids with
    more threads than a line has columns (several stretches per line) and
    with fewer (a 600 x 600 x 10 grid's line of 360,000 rows against 262,144
    threads), odd line counts and n not a multiple of the line; and the
    32-bit buffer predicate."""
    src = tmp_path / "part.cpp"
    src.write_text(r'''
#include <cstdio>
#include <vector>
#include "line_part.hpp"
static int check(uint32_t n, uint32_t line, uint32_t threads) {
  std::vector<unsigned char> seen(n, 0);
  const LinePart p = line_part(n, line, threads);
  for (uint32_t t = 0; t < threads; ++t)
    for (uint32_t q = t; q < p.items; q += threads) {
      uint32_t c, kb, ke;
      line_item(p, q, c, kb, ke);
      for (uint32_t k = kb; k < ke; ++k) {
        const uint64_t i = uint64_t(c) + uint64_t(k) * line;
        if (i < n && seen[i]++) { printf("twice %u %u %u row %llu\n", n, line, threads, (unsigned long long)i); return 1; }
      }
    }
  for (uint32_t i = 0; i < n; ++i)
    if (!seen[i]) { printf("missed %u %u %u row %u\n", n, line, threads, i); return 1; }
  return 0;
}
int main() {
  const uint32_t cases[][3] = {
    {64u * 64 * 16, 4096, 1024 * 256}, {64u * 64 * 16, 4096, 16 * 256},
    {600u * 600 * 10, 360000, 1024 * 256}, {640u * 640 * 16, 409600, 1024 * 256},
    {4096u * 9 + 77, 4096, 256}, {4096u * 9 + 77, 4096, 3 * 256}, {5000u * 15, 5000, 1024 * 256},
    {5000u * 15 - 1, 5000, 7 * 256}, {160u * 160 * 160, 25600, 1024 * 256}, {300001u, 37500, 1000 * 256},
  };
  int bad = 0;
  for (const auto& c : cases) bad |= check(c[0], c[1], c[2]);
  bad |= !dia_buf_fits(7, 153391689u, 4) ? 0 : 1;       // 7 n 4 bytes passes 2^32 - 16
  bad |= dia_buf_fits(7, 153391688u, 4) ? 0 : 1;        // just fits
  bad |= dia_buf_fits(4, 4096000u, 8) ? 0 : 1;
  printf(bad ? "FAIL\n" : "OK\n");
  return bad;
}''')
    exe = tmp_path / "part"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "clsparse_amd", "csrc"),
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout
