"""Learned phase-count / log-K initialiser: the MLP kernels against the CPU oracle
(oracle/mlp_oracle.py), and the hinted flash against the unhinted flash.

CPU tests pin the oracle itself (finite differences, softmax cross-entropy closed
forms, a planted mistake).  GPU tests compare the HIP loss/gradient (FP32) and the
BF16 inference with the oracle, train the network on flash results of the GPU
flash, and check that hints never change a converged phase label.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import mlp_oracle as mo  # noqa: E402

from nonenone_amd import fluids  # noqa: E402
from nonenone_amd.learned import param_count  # noqa: E402
from nonenone_amd.workloads import batched_flash_states  # noqa: E402

GRAD_RTOL = 2e-4   # FP32 kernel vs float64 oracle, relative to the gradient's max-abs
PRED_ATOL = 2e-2   # BF16 inference vs the oracle's BF16 emulation (probabilities, log-K)


def _batch(nc, n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    P, T, z = batched_flash_states(n, fluids.six_component("PR") if nc == 6 else
                                   fluids.h2o_co2_nc10("PR"), seed=seed)
    X = mo.features(P, T, z)
    label = rng.integers(0, 3, n)
    logk = rng.normal(0.0, 2.0, (n, nc))
    mask = (rng.random(n) < 0.6).astype(np.uint8)
    return P, T, z, X, label, logk, mask


# ---------------------------------------------------------------- CPU: oracle pins

def test_param_count_matches_oracle():
    for nc in (3, 4, 6):
        assert param_count(nc) == mo.param_count(nc)


@pytest.mark.parametrize("nc", [3, 6])
def test_oracle_gradient_matches_central_differences(nc):
    _, _, _, X, label, logk, mask = _batch(nc, 24, seed=nc)
    params = mo.init_params(nc, seed=1)
    _, g = mo.loss_grad(params, X, label, logk, mask, 0.7, nc)
    rng = np.random.Generator(np.random.PCG64(5))
    idx = rng.choice(params.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(params)
        e[i] = h
        lp, _ = mo.loss_grad(params + e, X, label, logk, mask, 0.7, nc)
        lm, _ = mo.loss_grad(params - e, X, label, logk, mask, 0.7, nc)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_oracle_softmax_closed_forms():
    nc = 3
    params = np.zeros(mo.param_count(nc))   # all outputs zero: equal logits
    X = np.random.default_rng(0).normal(size=(10, nc + 2))
    label = np.arange(10) % 3
    loss, g = mo.loss_grad(params, X, label, np.zeros((10, nc)), np.zeros(10), 1.0, nc)
    assert loss == pytest.approx(np.log(3.0), rel=1e-15)
    # output-layer bias gradient of the 3 logits = mean(softmax - onehot): sums to zero
    gb = g[-(3 + nc):][:3]
    assert abs(gb.sum()) < 1e-15


def test_planted_silu_mistake_is_detected():
    nc = 3
    _, _, _, X, label, logk, mask = _batch(nc, 16, seed=9)
    params = mo.init_params(nc, seed=2)
    _, g_bad = mo.loss_grad(params, X, label, logk, mask, 1.0, nc, dsilu=mo.sigmoid)
    i = 10  # a first-layer weight: its gradient passes through every SiLU derivative
    h = 1e-6
    e = np.zeros_like(params)
    e[i] = h
    fd = (mo.loss_grad(params + e, X, label, logk, mask, 1.0, nc)[0] -
          mo.loss_grad(params - e, X, label, logk, mask, 1.0, nc)[0]) / (2 * h)
    assert abs(fd - g_bad[i]) > 1e-5 * abs(fd) + 1e-7


def test_bf16_rounding_ties_to_even():
    vals = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, 1.0 + 2**-9, -2.5], np.float32)
    r = mo.bf16(vals)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2**-6) and r[3] == 1.0
    assert r[4] == np.float32(-2.5)


# ---------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("nc", [3, 6])
def test_gpu_loss_and_gradient_match_oracle(nc):
    import torch
    from nonenone_amd.learned import PhaseInitModel
    P, T, z, X, label, logk, mask = _batch(nc, 1000, seed=20 + nc)
    m = PhaseInitModel(nc, seed=3)
    params = m.get_params().cpu().numpy().astype(np.float64)
    Xg = m.features(P, T, z)
    assert np.allclose(Xg.cpu().numpy(), X, rtol=1e-6, atol=1e-6)
    loss, grad = m.loss_grad(Xg, label, logk, mask, logk_weight=0.7)
    ref_loss, ref_grad = mo.loss_grad(params, Xg.cpu().numpy().astype(np.float64), label,
                                      logk, mask, 0.7, nc)
    torch.cuda.synchronize()
    assert float(loss.item()) == pytest.approx(ref_loss, rel=1e-5)
    g = grad.cpu().numpy().astype(np.float64)
    scale = np.abs(ref_grad).max()
    assert np.abs(g - ref_grad).max() <= GRAD_RTOL * scale, np.abs(g - ref_grad).max() / scale


@pytest.mark.gpu
def test_gpu_adam_step_matches_oracle_update():
    import torch
    from nonenone_amd.learned import PhaseInitModel
    nc = 6
    P, T, z, X, label, logk, mask = _batch(nc, 777, seed=31)
    m = PhaseInitModel(nc, seed=4)
    p0 = m.get_params().cpu().numpy().astype(np.float64)
    Xg = m.features(P, T, z)
    lr = 1e-3
    m.train_step(Xg, label, logk, mask, lr, logk_weight=1.0)
    p1 = m.get_params().cpu().numpy().astype(np.float64)
    _, g = mo.loss_grad(p0, Xg.cpu().numpy().astype(np.float64), label, logk, mask, 1.0, nc)
    # first Adam step: m/(1-b1) = g, v/(1-b2) = g^2  ->  step = lr * g / (|g| + eps)
    expect = p0 - lr * g / (np.abs(g) + 1e-8)
    big = np.abs(g) > 1e-4 * np.abs(g).max()
    assert np.allclose(p1[big], expect[big], rtol=0, atol=lr * 1e-2)


@pytest.mark.gpu
@pytest.mark.parametrize("nc", [3, 6])
def test_gpu_bf16_inference_matches_oracle(nc):
    from nonenone_amd.learned import PhaseInitModel
    P, T, z, X, _, _, _ = _batch(nc, 3000, seed=40 + nc)
    m = PhaseInitModel(nc, seed=5)
    params = m.get_params().cpu().numpy()
    prob, logk = m.predict(m.features(P, T, z))
    rp, rl = mo.predict_bf16(params, X.astype(np.float32), nc)
    assert np.abs(prob.cpu().numpy() - rp).max() <= PRED_ATOL
    assert np.abs(logk.cpu().numpy() - rl).max() <= PRED_ATOL * max(1.0, np.abs(rl).max())


@pytest.mark.gpu
def test_gpu_hint_none_is_bit_identical_to_plain_flash():
    import torch
    from nonenone_amd.flash import ThreePhaseFlash
    fl = ThreePhaseFlash(fluids.six_component("PR"))
    P, T, z = batched_flash_states(4096, fl.fluid, seed=77, near_critical_fraction=0.1)
    a = fl.flash(P, T, z)
    b = fl.flash_seeded(P, T, z, np.zeros(len(P), np.uint8))
    torch.cuda.synchronize()
    for name in ("presence", "status", "beta", "x", "Z", "iterations"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name


@pytest.mark.gpu
def test_gpu_trained_hints():
    """Train on flash results of the GPU flash (nc=6 PR, the batched-flash
    distribution), then hint a fresh sample.  Order-only hints (the predicted
    phase count orders the batch) must give results bit-identical to the
    unhinted flash.  The path-changing single-phase hint, with its threshold
    calibrated on a validation sample, is measured: it must hint a fair share
    of the states and change at most 1 in 10^3 labels (it is not label-safe, so
    it is not the default)."""
    import torch
    from nonenone_amd import learned
    from nonenone_amd._abi import NN_HINT_ORDER, NN_HINT_SINGLE_PHASE
    from nonenone_amd.flash import ThreePhaseFlash
    fl = ThreePhaseFlash(fluids.six_component("PR"))
    model = learned.PhaseInitModel(6, seed=11)
    data = learned.generate(fl, model, 1 << 17, seed=1000)
    hist = learned.train(model, data, epochs=30, batch=4096, lr=3e-3)
    assert hist[-1] < hist[0]
    acc = learned.accuracy(model, data)
    majority = float(torch.bincount(data.label.to(torch.int64)).max().item()) / data.label.numel()
    assert acc > majority + 0.05, (acc, majority)
    P, T, z = batched_flash_states(1 << 14, fl.fluid, seed=5000, near_critical_fraction=0.1)
    cmp, _, _ = learned.evaluate(fl, model, P, T, z, modes=NN_HINT_ORDER)
    print("order-only", cmp)
    assert cmp["bit_identical"] and cmp["label_mismatches"] == 0, cmp
    Pv, Tv, zv = batched_flash_states(1 << 14, fl.fluid, seed=4000, near_critical_fraction=0.1)
    modes = NN_HINT_ORDER | NN_HINT_SINGLE_PHASE
    th, records = learned.calibrate(fl, model, Pv, Tv, zv, modes)
    print("single-phase hint: calibrated threshold", th,
          {k: (v["label_mismatches"], v["hinted_single"]) for k, v in records.items()})
    cmp, _, _ = learned.evaluate(fl, model, P, T, z, threshold=th or 0.9999, modes=modes)
    print("single-phase hint", cmp)
    assert cmp["hinted_single"] > 0.01 * len(P), cmp
    assert cmp["label_mismatches"] <= 1e-3 * cmp["hinted_single"], cmp
