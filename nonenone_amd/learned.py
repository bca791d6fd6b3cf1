"""Learned phase-count / log-K initialiser of the flash (no counterpart in the reference).

A small MLP — features (scaled ln P, scaled T, z), three hidden layers of 64 SiLU
units, outputs 3 phase-count logits + nc values of ln K_go — is trained on flash
results that this library's own GPU flash produces, with the library's own HIP
forward/backward/Adam kernels (FP32), and runs BF16 inference.  Its prediction
becomes a per-state hint of ``ThreePhaseFlash.flash_seeded``:

* NN_HINT_ORDER: the predicted phase count orders the batch (states of the
  same predicted count share wavefronts); the solver path is the reference's,
  so results are bit-identical to the unhinted flash — the mode that can never
  change a label;
* NN_HINT_SINGLE_PHASE (p(1 phase) >= threshold): the flash skips its first
  gas-oil two-phase start and goes straight to the single-phase candidate;
* NN_HINT_LOG_K (p(2) or p(3) >= threshold): that start uses exp(predicted ln K)
  instead of Wilson's K;
* otherwise the reference path runs unchanged.

With the two path-changing hints, the active-set equilibration's stability
tests, the three-phase flash and the recovery chain still run and judge the
hinted start, but that judgement is not complete: the reference's stability
seeds can miss a split that its Wilson-started two-phase flash finds, so a
skipped start can change a label (measured: about 3 in 10^4 hinted states).
``calibrate`` lets the full method judge the predictor on a validation sample
and keeps the lowest confidence threshold with zero changes there;
``evaluate`` measures what hints change on fresh states.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi, workloads
from .flash import FlashBatchResult, ThreePhaseFlash, _dev, _ptr, _stream

HIDDEN = 64
CLASSES = 3


def param_count(nc: int) -> int:
    nin, nout = nc + 2, CLASSES + nc
    return HIDDEN * nin + HIDDEN + 2 * (HIDDEN * HIDDEN + HIDDEN) + nout * HIDDEN + nout


class PhaseInitModel:
    """Owns an nn_mlp (parameters and Adam moments on the current HIP device)."""

    def __init__(self, nc: int, seed: int = 0):
        self.lib = _abi.lib()
        if not torch.cuda.is_available():
            raise RuntimeError("no HIP device: this library runs on the GPU only")
        self.nc = nc
        h = ctypes.c_void_p()
        _abi.check(self.lib.nn_mlp_create(nc, ctypes.c_uint64(seed), ctypes.byref(h)),
                   "nn_mlp_create")
        self.handle = h
        self.nparam = int(self.lib.nn_mlp_param_count(self.handle))
        self._loss = torch.zeros(1, dtype=torch.float32, device="cuda")

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.nn_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters ---------------------------------------------------------
    def get_params(self) -> torch.Tensor:
        out = torch.empty(self.nparam, dtype=torch.float32, device="cuda")
        _abi.check(self.lib.nn_mlp_get_params(self.handle, _ptr(out), _stream()),
                   "nn_mlp_get_params")
        return out

    def set_params(self, params) -> None:
        p = _dev(params, torch.float32)
        if p.numel() != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {p.numel()}")
        _abi.check(self.lib.nn_mlp_set_params(self.handle, _ptr(p), _stream()),
                   "nn_mlp_set_params")
        torch.cuda.current_stream().synchronize()

    def reset_optimizer(self) -> None:
        _abi.check(self.lib.nn_mlp_reset_optimizer(self.handle), "nn_mlp_reset_optimizer")

    # ---- kernels --------------------------------------------------------------
    def features(self, P, T, z) -> torch.Tensor:
        P, T, z = _dev(P, torch.float64), _dev(T, torch.float64), _dev(z, torch.float64)
        n = P.shape[0]
        X = torch.empty(n, self.nc + 2, dtype=torch.float32, device=P.device)
        _abi.check(self.lib.nn_mlp_features(self.handle, n, _ptr(P), _ptr(T), _ptr(z), _ptr(X),
                                            _stream()), "nn_mlp_features")
        return X

    def loss_grad(self, X, label, logk, mask, logk_weight: float = 1.0):
        """(loss, gradient) of the batch, no update."""
        X, label, logk, mask = self._batch(X, label, logk, mask)
        grad = torch.empty(self.nparam, dtype=torch.float32, device="cuda")
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        _abi.check(self.lib.nn_mlp_loss_grad(self.handle, X.shape[0], _ptr(X), _ptr(label),
                                             _ptr(logk), _ptr(mask), float(logk_weight),
                                             _ptr(loss), _ptr(grad), _stream()),
                   "nn_mlp_loss_grad")
        return loss, grad

    def train_step(self, X, label, logk, mask, lr: float, logk_weight: float = 1.0):
        """One Adam step; returns the (device) loss before the step."""
        X, label, logk, mask = self._batch(X, label, logk, mask)
        _abi.check(self.lib.nn_mlp_train_step(self.handle, X.shape[0], _ptr(X), _ptr(label),
                                              _ptr(logk), _ptr(mask), float(logk_weight),
                                              float(lr), _ptr(self._loss), _stream()),
                   "nn_mlp_train_step")
        return self._loss

    def predict(self, X):
        """BF16 inference -> (class probabilities n x 3, log-K n x nc)."""
        X = _dev(X, torch.float32)
        n = X.shape[0]
        prob = torch.empty(n, CLASSES, dtype=torch.float32, device=X.device)
        logk = torch.empty(n, self.nc, dtype=torch.float32, device=X.device)
        _abi.check(self.lib.nn_mlp_predict(self.handle, n, _ptr(X), _ptr(prob), _ptr(logk),
                                           _stream()), "nn_mlp_predict")
        return prob, logk

    def hints(self, prob, logk, threshold: float = 0.95, modes: int = _abi.NN_HINT_ORDER):
        n = prob.shape[0]
        hint = torch.empty(n, dtype=torch.uint8, device=prob.device)
        seed = torch.empty(n, self.nc, dtype=torch.float64, device=prob.device)
        _abi.check(self.lib.nn_mlp_hints(self.handle, n, _ptr(prob), _ptr(logk),
                                         float(threshold), int(modes), _ptr(hint), _ptr(seed),
                                         _stream()), "nn_mlp_hints")
        return hint, seed

    def _batch(self, X, label, logk, mask):
        X = _dev(X, torch.float32)
        label = _dev(label, torch.int32)
        logk = _dev(logk, torch.float32)
        mask = _dev(mask, torch.uint8)
        n = X.shape[0]
        if not (label.shape[0] == n and logk.shape == (n, self.nc) and mask.shape[0] == n):
            raise ValueError("inconsistent batch shapes")
        return X, label, logk, mask


# ---- training data from the method itself ------------------------------------

@dataclass
class TrainingSet:
    X: torch.Tensor       # n x (nc+2) features
    label: torch.Tensor   # n, phase count - 1
    logk: torch.Tensor    # n x nc, ln(x_gas / x_oil) where both slots are present
    mask: torch.Tensor    # n, 1 where logk applies


def targets(res: FlashBatchResult, nc: int):
    """Labels and ln K_go targets from converged flash results (device tensors);
    returns (keep, label, logk, mask) with keep = converged states."""
    keep = res.status == _abi.NN_STATUS_OK
    pres = res.presence.to(torch.int32)
    count = (pres & 1) + ((pres >> 1) & 1) + ((pres >> 2) & 1)
    label = (count - 1).clamp(0, 2)
    og = ((pres & 3) == 3)
    xo = res.x[:, 0, :].clamp_min(1e-300)
    xg = res.x[:, 1, :].clamp_min(1e-300)
    logk = torch.where(og[:, None], torch.log(xg) - torch.log(xo), torch.zeros_like(xo))
    keep = keep & (count > 0)
    return keep, label, logk.to(torch.float32), og.to(torch.uint8)


def generate(flash: ThreePhaseFlash, model: PhaseInitModel, n: int, seed: int,
             near_critical_fraction: float = 0.1, chunk: int = 1 << 20) -> TrainingSet:
    """Flash n seeded states of the batched-flash distribution on the GPU and keep the
    converged ones as a training set."""
    parts = []
    for a in range(0, n, chunk):
        m = min(chunk, n - a)
        P, T, z = workloads.batched_flash_states(m, flash.fluid, seed=seed + a,
                                                 near_critical_fraction=near_critical_fraction)
        res = flash.flash(P, T, z)
        X = model.features(P, T, z)
        keep, label, logk, mask = targets(res, flash.nc)
        parts.append((X[keep], label[keep], logk[keep], mask[keep]))
    cat = [torch.cat([p[i] for p in parts]) for i in range(4)]
    return TrainingSet(*cat)


def train(model: PhaseInitModel, data: TrainingSet, epochs: int = 8, batch: int = 16384,
          lr: float = 2e-3, logk_weight: float = 1.0, seed: int = 0, log=None) -> list[float]:
    """Mini-batch Adam on the GPU; returns the mean loss of every epoch."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    n = data.X.shape[0]
    history = []
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda", generator=g)
        rate = lr * (0.5 * (1.0 + np.cos(np.pi * ep / max(epochs, 1))))  # cosine decay
        total = torch.zeros(1, dtype=torch.float32, device="cuda")
        steps = 0
        for a in range(0, n, batch):
            idx = perm[a:a + batch]
            loss = model.train_step(data.X[idx], data.label[idx], data.logk[idx],
                                    data.mask[idx], rate, logk_weight)
            total += loss
            steps += 1
        history.append(float(total.item()) / max(steps, 1))
        if log:
            log(f"epoch {ep}: mean loss {history[-1]:.4f} (lr {rate:.2e})")
    return history


def accuracy(model: PhaseInitModel, data: TrainingSet) -> float:
    prob, _ = model.predict(data.X)
    return float((prob.argmax(1).to(torch.int32) == data.label).float().mean().item())


# ---- what the hints change ----------------------------------------------------

def compare(plain: FlashBatchResult, seeded: FlashBatchResult, rel_floor: float = 1e-12):
    """Label mismatches and the largest relative value difference between two flash
    results of the same states (converged states of the plain flash)."""
    ok = (plain.status == 0)
    labels = int(((plain.presence != seeded.presence) & ok).sum().item())
    status = int((plain.status != seeded.status).sum().item())
    worst = 0.0
    for name in ("beta", "x", "Z"):
        a = getattr(plain, name)[ok]
        b = getattr(seeded, name)[ok]
        scale = torch.maximum(a.abs(), b.abs()).clamp_min(rel_floor)
        if a.numel():
            worst = max(worst, float(((a - b).abs() / scale).max().item()))
    return dict(label_mismatches=labels, status_mismatches=status, max_rel_diff=worst)


CALIBRATION_THRESHOLDS = (0.95, 0.98, 0.99, 0.995, 0.999, 0.9999)


def calibrate(flash: ThreePhaseFlash, model: PhaseInitModel, P, T, z, modes: int,
              thresholds=CALIBRATION_THRESHOLDS):
    """The lowest confidence threshold at which the hinted flash reproduces the
    unhinted flash's labels and statuses on a validation sample (the full method
    judges the predictor), or None when no threshold does.  Returns (threshold,
    per-threshold comparison records)."""
    records = {}
    for th in thresholds:
        cmp, _, _ = evaluate(flash, model, P, T, z, threshold=th, modes=modes)
        records[th] = cmp
        if cmp["label_mismatches"] == 0 and cmp["status_mismatches"] == 0:
            return th, records
    return None, records


def evaluate(flash: ThreePhaseFlash, model: PhaseInitModel, P, T, z, threshold: float = 0.95,
             modes: int = _abi.NN_HINT_ORDER, repeats: int = 1):
    """Plain flash vs hinted flash on the same states: labels, values, times, hint counts."""
    P, T, z = _dev(P, torch.float64), _dev(T, torch.float64), _dev(z, torch.float64)
    n = P.shape[0]
    plain = FlashBatchResult.empty(n, flash.nc)
    seeded = FlashBatchResult.empty(n, flash.nc)
    torch.cuda.synchronize()

    def timed(fn):
        best = float("inf")
        for _ in range(repeats):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    t_plain = timed(lambda: flash.flash(P, T, z, out=plain))
    state = {}

    def hinted():
        X = model.features(P, T, z)
        prob, logk = model.predict(X)
        hint, seedk = model.hints(prob, logk, threshold, modes)
        state["hint"] = hint
        flash.flash_seeded(P, T, z, hint, seedk, out=seeded)

    t_seeded = timed(hinted)
    hint = state["hint"]
    path = hint & 3
    cmp = compare(plain, seeded)
    identical = all(torch.equal(getattr(plain, f), getattr(seeded, f))
                    for f in ("presence", "status", "converged", "iterations", "beta",
                              "saturation", "x", "Z", "rho_m", "k_go", "k_wo"))
    cmp.update(
        n=n, plain_s=t_plain, seeded_s=t_seeded, speedup=t_plain / t_seeded,
        bit_identical=identical,
        hinted_single=int((path == _abi.NN_HINT_SINGLE_PHASE).sum().item()),
        hinted_log_k=int((path == _abi.NN_HINT_LOG_K).sum().item()),
        predicted_counts=torch.bincount((hint >> 2).to(torch.int64) & 3,
                                        minlength=3).tolist(),
        iterations_plain=int(plain.iterations.to(torch.int64).sum().item()),
        iterations_seeded=int(seeded.iterations.to(torch.int64).sum().item()))
    return cmp, plain, seeded
