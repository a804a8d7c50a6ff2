"""Training, storage and evaluation of the learned FSAI network.

The network and the factor live in the HIP library (clsparseAmdMlp*,
clsparseAmdFsai*); this module drives them: the seeded family of 2-D
anisotropic diffusion matrices the network is trained on, the training loop
(Hutchinson loss, Adam, all on the GPU), and weight files (.npy of the FP32
parameter vector plus a JSON record with its SHA-256 and how it was made).
"""
from __future__ import annotations

import concurrent.futures as cf
import hashlib
import json
import math
import os
import time

import numpy as np

from . import gen_aniso2d, gen_diffusion3d
from .device import Control, CsrMatrix, Fsai, Mlp

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
DEFAULT_WEIGHTS = os.path.join(WEIGHTS_DIR, "fsai_aniso2d.npy")


# ---- the training family ------------------------------------------------------

def aniso2d_params(i: int, base_seed: int = 20240) -> dict:
    """Matrix i of the family: log10(ratio) ~ U[0, 3], theta ~ U[0, pi),
    log-coefficient GRF of correlation length 16 and sigma 1."""
    rng = np.random.default_rng(base_seed + 7919 * i)
    return {"seed": base_seed + i, "ratio": float(10.0 ** rng.uniform(0.0, 3.0)),
            "theta": float(rng.uniform(0.0, math.pi)), "corr_len": 16.0, "sigma": 1.0}


def gen_family(i: int, grid: int = 512):
    p = aniso2d_params(i)
    return gen_aniso2d(grid, grid, corr_len=p["corr_len"], sigma=p["sigma"], ratio=p["ratio"],
                       theta=p["theta"], seed=p["seed"])


# ---- weights ------------------------------------------------------------------

def save_weights(path: str, params: np.ndarray, record: dict) -> str:
    params = np.ascontiguousarray(params, dtype=np.float32)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, params, allow_pickle=False)
    digest = hashlib.sha256(params.tobytes()).hexdigest()
    rec = dict(record)
    rec.update({"sha256_float32_le": digest, "num_params": int(params.size)})
    with open(os.path.splitext(path)[0] + ".json", "w") as f:
        json.dump(rec, f, indent=1, sort_keys=True)
    return digest


def load_weights(path: str = DEFAULT_WEIGHTS) -> tuple[np.ndarray, dict]:
    """Parameters and their record; raises if the checksum does not match."""
    params = np.load(path, allow_pickle=False).astype(np.float32)
    with open(os.path.splitext(path)[0] + ".json") as f:
        rec = json.load(f)
    digest = hashlib.sha256(np.ascontiguousarray(params).tobytes()).hexdigest()
    if digest != rec.get("sha256_float32_le"):
        raise ValueError(f"{path}: checksum mismatch")
    if params.size != Mlp.num_params():
        raise ValueError(f"{path}: {params.size} parameters, library expects {Mlp.num_params()}")
    return params, rec


# ---- training -------------------------------------------------------------------

class _Slot:
    """One device matrix of the family's fixed pattern whose values are swapped."""

    def __init__(self, ctl: Control, rp, ci, v, m):
        import torch
        self.A = CsrMatrix(rp, ci, v, m)
        self.A.meta_create(ctl)
        self.f = Fsai(self.A, ctl)
        self.torch = torch

    def load(self, v, ctl: Control) -> None:
        self.A.va.copy_(self.torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)))
        self.f.update_values(self.A, ctl)

    def release(self) -> None:
        self.f.release()
        self.A.meta_delete()


def train(net: Mlp, ctl: Control, train_ids, steps_per_matrix: int = 4, lr: float = 2e-3,
          lr_min: float = 1e-4, probes: int = 8, grid: int = 512, log=None,
          time_budget_s: float | None = None) -> dict:
    """Trains on the family members `train_ids` (one pass), generating the
    next matrix on a host thread while the GPU trains on the current one.
    Cosine learning-rate decay from lr to lr_min over all steps."""
    ids = list(train_ids)
    total = len(ids) * steps_per_matrix
    slot = None
    losses = []
    step = 0
    t0 = time.perf_counter()
    gpu_s = 0.0
    with cf.ThreadPoolExecutor(max_workers=1) as pool:
        fut = pool.submit(gen_family, ids[0], grid) if ids else None
        for k, i in enumerate(ids):
            rp, ci, v, m = fut.result()
            if k + 1 < len(ids):
                fut = pool.submit(gen_family, ids[k + 1], grid)
            if slot is None:
                slot = _Slot(ctl, rp, ci, v, m)
            else:
                slot.load(v, ctl)
            g0 = time.perf_counter()
            for s in range(steps_per_matrix):
                frac = step / max(total - 1, 1)
                cur = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * frac))
                loss, _ = slot.f.train_step(net, slot.A, ctl, probes=probes,
                                            seed=1_000_003 * i + s, lr=cur)
                losses.append(loss)
                step += 1
            gpu_s += time.perf_counter() - g0
            if log and (k % 100 == 0 or k + 1 == len(ids)):
                log(f"train matrix {k + 1}/{len(ids)} id {i} loss {losses[-1]:.4f} "
                    f"({time.perf_counter() - t0:.0f}s)")
            if time_budget_s is not None and time.perf_counter() - t0 > time_budget_s:
                if log:
                    log(f"time budget reached after {k + 1} matrices")
                ids = ids[:k + 1]
                break
    if slot is not None:
        slot.release()
    n_first = max(1, len(losses) // 20)
    return {"matrices": len(ids), "steps": step, "probes": probes, "lr": lr, "lr_min": lr_min,
            "loss_first": float(np.mean(losses[:n_first])) if losses else None,
            "loss_last": float(np.mean(losses[-n_first:])) if losses else None,
            "seconds": round(time.perf_counter() - t0, 2),
            "train_step_seconds": round(gpu_s, 5)}


def evaluate(net: Mlp | None, ctl: Control, ids, probes: int = 8, grid: int = 512) -> list:
    """Hutchinson loss of each family member (no update); net None = Jacobi."""
    out = []
    slot = None
    zero = Mlp(seed=0) if net is None else None  # zero output layer: exactly Jacobi
    use = net if net is not None else zero
    try:
        for i in ids:
            rp, ci, v, m = gen_family(i, grid)
            if slot is None:
                slot = _Slot(ctl, rp, ci, v, m)
            else:
                slot.load(v, ctl)
            loss, _ = slot.f.train_step(use, slot.A, ctl, probes=probes, seed=97 * i + 5, lr=0.0)
            out.append(loss)
    finally:
        if slot is not None:
            slot.release()
        if zero is not None:
            zero.release()
    return out


def gen_target3d(n1d: int = 160, anisotropy: float = 100.0, sigma: float = 1.0, seed: int = 31):
    """The held-out 3-D problem: 7-point variable-coefficient diffusion,
    kappa ~ LogNormal(0, sigma), z faces scaled by `anisotropy`."""
    return gen_diffusion3d(n1d, n1d, n1d, sigma=sigma, ax=1.0, ay=1.0, az=anisotropy, seed=seed)
