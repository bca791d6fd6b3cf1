"""nn_flash_batch_static (the allocation-free flash: work queue and deferred list in
the library's static device scratch) against nn_flash_batch on the same states,
bit for bit, on two calls in a row, the second with a smaller batch; and its
refusal of a batch beyond NN_STATIC_MAX_STATES before any device work."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from nonenone_amd import _abi, fluids, workloads
from nonenone_amd.flash import FlashBatchResult, ThreePhaseFlash, _dev, _ptr, _stream

FIELDS = ("presence", "converged", "status", "iterations", "beta", "saturation", "x", "Z",
          "rho_m", "k_go", "k_wo")


def _static(fl, P, T, z):
    P, T, z = _dev(P, torch.float64), _dev(T, torch.float64), _dev(z, torch.float64)
    n = P.shape[0]
    out = FlashBatchResult.empty(n, fl.nc, P.device)
    for name in FIELDS:                      # stale values must not survive
        getattr(out, name).fill_(77)
    abi = out.abi()
    _abi.check(fl.lib.nn_flash_batch_static(fl.handle, n, _ptr(P), _ptr(T), _ptr(z),
                                            ctypes.byref(abi), _stream()),
               "nn_flash_batch_static")
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["PR", "SW"])
def test_gpu_static_flash_matches_flash_batch(model):
    fl = ThreePhaseFlash(fluids.h2o_co2_nc10(model))
    for call, (n, seed) in enumerate([(3000, 11), (700, 12)]):   # second call smaller
        print(f"static flash {model}: call {call}, n={n}, seed={seed}", flush=True)
        P, T, z = workloads.batched_flash_states(n, fl.fluid, seed=seed, near_critical_fraction=0.1)
        ref = fl.flash(P, T, z)
        got = _static(fl, P, T, z)
        torch.cuda.synchronize()
        for name in FIELDS:
            a, b = getattr(got, name), getattr(ref, name)
            assert torch.equal(a, b), f"call {call}: {name}"
        if model == "SW":
            print("states not converged / deferred-path statuses:",
                  np.bincount(ref.status.cpu().numpy(), minlength=5))


@pytest.mark.gpu
def test_gpu_static_flash_refuses_oversized_batch():
    fl = ThreePhaseFlash(fluids.h2o_co2_nc10("PR"))
    n = (1 << 20) + 1
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")
    out = FlashBatchResult.empty(1, fl.nc, "cuda")
    abi = out.abi()
    rc = fl.lib.nn_flash_batch_static(fl.handle, n, _ptr(dummy), _ptr(dummy), _ptr(dummy),
                                      ctypes.byref(abi), _stream())
    assert rc == _abi.NN_E_INVALID_ARGUMENT
    assert b"static scratch" in fl.lib.nn_last_error()
