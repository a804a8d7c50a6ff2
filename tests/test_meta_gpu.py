"""clsparseCsrMetaSize / Create / Delete through the device path, bit-exact with
the oracle's restatement of csr-meta.hpp (itself pinned by hand-derived known
answers in test_oracle.py)."""
import json
import os

import numpy as np
import pytest
import torch

import clsparse_amd as cls
import oracle
from gpu_helpers import rp_from_lengths, say

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "row_blocks_known.json")


@pytest.fixture(scope="module")
def ctl():
    c = cls.Control()
    yield c
    c.release()


def _meta_roundtrip(ctl, rp, ncols):
    nnz = int(rp[-1])
    A = cls.CsrMatrix(rp, np.zeros(nnz, np.uint32), np.zeros(nnz), max(ncols, 1))
    before = A.meta_size(ctl)
    A.meta_create(ctl)
    after = A.meta_size(ctl)
    entries = A.row_blocks()
    A.meta_delete()
    return before, after, entries


@pytest.mark.parametrize("case", json.load(open(GOLDEN))["cases"], ids=lambda c: c["name"])
def test_known_answers_through_device(ctl, case):
    rp = rp_from_lengths(case["row_lengths"])
    before, after, entries = _meta_roundtrip(ctl, rp, 1)
    assert [int(e) for e in entries] == case["entries"]
    assert after == case["row_block_size"]
    assert before == case["size_pass"]


def test_workload_patterns_bit_exact(ctl):
    pats = {
        "poisson256": cls.gen_poisson2d(256, 256)[0],
        "diffusion128": cls.gen_diffusion3d(128, 128, 128)[0],
        "powerlaw2^21": cls.gen_powerlaw(1 << 21, 16.0, 1.5, 200000, 8, 100000, seed=1)[0],
    }
    for name, rp in pats.items():
        say(f"meta {name}")
        before, after, entries = _meta_roundtrip(ctl, rp, rp.size - 1)
        ref, size = oracle.row_blocks(rp)
        assert after == size and before == oracle.row_blocks(rp, count_only=True)
        assert np.array_equal(entries, ref), name


def test_recreate_and_delete(ctl):
    rp, ci, v, m = cls.gen_poisson2d(30, 30)
    A = cls.CsrMatrix(rp, ci, v, m)
    A.meta_create(ctl)
    first = A.s.meta
    A.meta_create(ctl)  # replaces (csr-meta.cpp L60-67)
    assert A.s.meta and np.array_equal(A.row_blocks(), oracle.row_blocks(rp)[0])
    del first
    A.meta_delete()
    assert not A.s.meta
    A.meta_delete()  # no meta: still success


def test_malformed_row_pointer_rejected(ctl):
    rp = np.array([0, 5, 3, 8], dtype=np.uint32)  # decreasing
    A = cls.CsrMatrix(rp, np.zeros(8, np.uint32), np.zeros(8), 4)
    with pytest.raises(cls.ClsparseError) as e:
        A.meta_create(ctl)
    assert e.value.status == cls.clsparseInvalidValue
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    del x
