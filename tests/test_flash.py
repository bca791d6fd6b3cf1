"""Flash / stability / EOS parity: HIP kernels against the reference.

Oracles, in order of strength:
  * the reference itself, compiled from its sources (oracle/_ref, when built);
  * golden fixtures in tests/golden/ written by tests/golden/make_golden.py
    from that build (always present);
  * the reference's own known answers (test/src/unit/natural_core_test.cpp:84-125)
    and physical invariants (test/src/unit/three_phase_flash_test.cpp:116-191).
Bar: phase labels / status bit-exact, values within 1e-10 relative.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from nonenone_amd import fluids
from nonenone_amd._abi import FlashOptions, FluidDesc, NN_MAX_NC, NN_MODEL_PR
from nonenone_amd.workloads import batched_flash_states

import flashcmp
import libm_conditioning
import refharness

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = {
    "flash_nc3_pr": lambda: fluids.h2o_co2_nc10("PR"),
    "flash_nc3_sw": lambda: fluids.h2o_co2_nc10("SW"),
    "flash_nc6_pr": lambda: fluids.six_component("PR"),
    "flash_nc6_sw": lambda: fluids.six_component("SW"),
    "flash_nc3_cpa": fluids.cpa_h2o_co2_nc10,
}


def core_test_fluid() -> FluidDesc:
    """The 6-component PR mixture of natural_core_test.cpp:58-82 (eos type 1)."""
    d = FluidDesc()
    d.nc = 6
    d.thermo_model = NN_MODEL_PR
    d.eos_type = 1
    d.water_component = -1
    d.omega_a, d.omega_b = 0.4572355, 0.0779691
    d.m1, d.m2 = 2.414213562373095, -0.414213562373095
    tc = [189.515, 304.2, 387.607, 597.497, 698.515, 875.0]
    pc = [4580011.59, 7386592.50, 4095515.97, 3345244.875, 1768374.5625, 1169006.79]
    vc = [9.97012032965401e-5, 9.26344713533338e-5, 0.000217076707259486,
          0.000381162235869935, 0.000721410148917871, 0.00113570073874421]
    om = [0.00854, 0.228, 0.16733, 0.38609, 0.80784, 1.23141]
    mw = [0.0161594, 0.04401, 0.0455725, 0.11774, 0.248827, 0.48152]
    k = [[0.0, 0.00070981, 0.00077754, 0.0100, 0.0110, 0.0110],
         [0.00070981, 0.0, 0.1500, 0.1500, 0.1500, 0.1500],
         [0.00077754, 0.1500, 0.0, 0.0, 0.0, 0.0],
         [0.0100, 0.1500, 0.0, 0.0, 0.0, 0.0],
         [0.0110, 0.1500, 0.0, 0.0, 0.0, 0.0],
         [0.0110, 0.1500, 0.0, 0.0, 0.0, 0.0]]
    for i in range(6):
        d.tc[i], d.pc[i], d.vc[i], d.acentric[i], d.molar_mass[i] = tc[i], pc[i], vc[i], om[i], mw[i]
        for j in range(6):
            d.kij[i * NN_MAX_NC + j] = k[i][j]
    return d


# natural_core_test.cpp:86-108
CORE_X = [0.3246914, 0.0128351, 0.2278401, 0.2606985, 0.1134144, 0.0605206]
CORE_Y = [0.808671, 0.025284, 0.1487798, 0.0175878, 0.0006759, 1.51e-06]
CORE_ZL, CORE_ZV = 0.66880987807546899, 0.84105046763526459
CORE_FL = [11148938.7289865, 285739.262764987, 958802.358789237, 32132.229549948,
           222.962965571357, 0.0303098653818211]
CORE_FV = [11364784.8024743, 289457.590456904, 964022.546010914, 31595.5016348791,
           198.090892606485, 0.0143840907404893]


def load_case(name):
    data = np.load(GOLDEN / f"{name}.npz")
    return {k: data[k] for k in data.files}


# ------------------------------------------------------------------ CPU tests
@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_fixture_invariants(name):
    """The committed fixtures satisfy the flash invariants of three_phase_flash_test.cpp:116-191."""
    g = load_case(name)
    conv = g["converged"] == 1
    assert conv.mean() > 0.95
    beta = g["beta"][conv]
    x = g["x"][conv]
    z = g["z"][conv]
    z = z / z.sum(axis=1, keepdims=True)
    assert np.all(np.abs(beta.sum(axis=1) - 1.0) < 1e-10)
    assert np.all(np.abs(g["saturation"][conv].sum(axis=1) - 1.0) < 1e-10)
    active = beta > 1e-10
    rec = np.einsum("np,npc->nc", np.where(active, beta, 0.0), x)
    assert np.max(np.abs(rec - z)) < 2e-8
    for p in range(3):
        on = active[:, p]
        assert np.all(np.abs(x[on, p].sum(axis=1) - 1.0) < 1e-10)
    # presence bits follow the phase fractions
    bits = (active * np.array([1, 2, 4])).sum(axis=1)
    assert np.array_equal(bits, g["presence"][conv])


@pytest.mark.skipif(not refharness.available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_fixture_matches_reference_build(name):
    g = load_case(name)
    ref = refharness.Reference(CASES[name](), FlashOptions.defaults())
    out = ref.flash(g["P"], g["T"], g["z"])
    for k in ("presence", "converged", "status", "iterations", "beta", "x", "Z"):
        assert np.array_equal(out[k], g[k]), k


@pytest.mark.skipif(not refharness.available(), reason="reference build (oracle/_ref) not present")
def test_reference_build_reproduces_core_test_goldens():
    ref = refharness.Reference(core_test_fluid(), FlashOptions.defaults())
    P = np.full(2, 15.0e6)
    T = np.full(2, 387.45)
    Z, lnphi, fug, st = ref.eos_phase(P, T, np.array([CORE_X, CORE_Y]), np.array([0, 1]))
    assert np.all(st == 0)
    assert abs(Z[0] - CORE_ZL) < 2e-12 and abs(Z[1] - CORE_ZV) < 2e-12
    np.testing.assert_allclose(fug[0], CORE_FL, rtol=2e-8)
    np.testing.assert_allclose(fug[1], CORE_FV, rtol=2e-8)


# ------------------------------------------------------------------ GPU tests
def _flash_gpu(fluid):
    from nonenone_amd.flash import ThreePhaseFlash
    return ThreePhaseFlash(fluid)


@pytest.mark.gpu
def test_gpu_eos_core_test_goldens():
    """E1-E5 against natural_core_test.cpp:96-108 (Z 2e-12, fugacities 2e-8)."""
    fl = _flash_gpu(core_test_fluid())
    P = np.full(2, 15.0e6)
    T = np.full(2, 387.45)
    Z, lnphi, fug, st = fl.eos_phase(P, T, np.array([CORE_X, CORE_Y]), np.array([0, 1]))
    Z, fug, st = Z.cpu().numpy(), fug.cpu().numpy(), st.cpu().numpy()
    assert np.all(st == 0)
    assert abs(Z[0] - CORE_ZL) < 2e-12
    assert abs(Z[1] - CORE_ZV) < 2e-12
    np.testing.assert_allclose(fug[0], CORE_FL, rtol=2e-8)
    np.testing.assert_allclose(fug[1], CORE_FV, rtol=2e-8)


def _gibbs(fluid, P, T, mine, ref):
    """Per-state Gibbs-energy difference of two flash answers (reference build's EOS),
    or None without the reference build."""
    if not refharness.available():
        return None
    return flashcmp.gibbs_diff(refharness.Reference(fluid, FlashOptions.defaults()), P, T, mine, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_flash_matches_golden(name):
    g = load_case(name)
    fl = _flash_gpu(CASES[name]())
    mine = flashcmp.to_numpy(fl.flash(g["P"], g["T"], g["z"]))
    n_status, n_label, worst, wi, sel = flashcmp.compare_flash(mine, g)
    unexplained, conditioned = libm_conditioning.split_value_failures(
        flashcmp.per_state_value_diff(mine, g), sel, flashcmp.REL_TOL, name,
        gibbs=_gibbs(CASES[name](), g["P"], g["T"], mine, g), beta=mine["beta"])
    print(f"{name}: status mismatches {n_status}, label mismatches {n_label}, "
          f"max rel diff {worst:.3e} (state {wi}), compared {int(sel.sum())}; above tolerance: "
          f"{len(conditioned)} libm-conditioned {conditioned}, {len(unexplained)} other {unexplained}")
    assert n_status == 0
    assert n_label == 0
    assert not unexplained


@pytest.mark.gpu
@pytest.mark.skipif(not refharness.available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_flash_matches_reference_build(name):
    fluid = CASES[name]()
    n = 1024 if "pr" in name else 256
    P, T, z = batched_flash_states(n, fluid, seed=7, near_critical_fraction=0.1)
    ref = refharness.Reference(fluid, FlashOptions.defaults()).flash(P, T, z, threads=8)
    mine = flashcmp.to_numpy(_flash_gpu(fluid).flash(P, T, z))
    n_status, n_label, worst, wi, sel = flashcmp.compare_flash(mine, ref)
    unexplained, conditioned = libm_conditioning.split_value_failures(
        flashcmp.per_state_value_diff(mine, ref), sel, flashcmp.REL_TOL,
        name.replace("flash_", "live_"), gibbs=_gibbs(fluid, P, T, mine, ref), beta=mine["beta"])
    print(f"{name}: n={n} status mismatches {n_status}, label mismatches {n_label}, "
          f"max rel diff {worst:.3e}; above tolerance: {len(conditioned)} libm-conditioned "
          f"{conditioned}, {len(unexplained)} other {unexplained}")
    assert n_status == 0 and n_label == 0
    assert not unexplained


@pytest.mark.gpu
def test_gpu_stability_matches_golden_single_phase():
    """stabilityTest of each converged single-phase golden state reports stable."""
    g = load_case("flash_nc3_pr")
    fl = _flash_gpu(fluids.h2o_co2_nc10("PR"))
    sel = (g["converged"] == 1) & np.isin(g["presence"], [1, 2, 4])
    r = fl.stability_test(g["P"][sel], g["T"][sel], g["z"][sel], g["presence"][sel], g["x"][sel])
    assert np.all(r["status"].cpu().numpy() == 0)
    assert np.all(r["valid"].cpu().numpy() == 1)
    assert np.all(r["stable"].cpu().numpy() == 1)


# ------------------------------------------------------------------ CPA (E6)
def _cpa_states(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    P = rng.uniform(1.0e6, 40.0e6, n)
    T = rng.uniform(300.0, 500.0, n)
    x = rng.dirichlet(np.ones(3), n)
    role = rng.integers(0, 3, n).astype(np.int32)
    return P, T, x, role


@pytest.mark.gpu
@pytest.mark.skipif(not refharness.available(), reason="reference build (oracle/_ref) not present")
def test_gpu_cpa_eos_matches_reference_build():
    """E6: role-aware CPA phase results (4C water + CO2 B2 cross-association,
    single-donor site solver) against the reference build."""
    fluid = fluids.cpa_h2o_co2_nc10()
    P, T, x, role = _cpa_states(512, 31)
    rZ, rlnphi, rfug, rst = refharness.Reference(fluid, FlashOptions.defaults()).eos_phase(P, T, x, role)
    Z, lnphi, fug, st = (t.cpu().numpy() for t in _flash_gpu(fluid).eos_phase(P, T, x, role))
    assert np.array_equal(st, rst)
    ok = rst == 0
    dz = np.abs(Z[ok] - rZ[ok]) / np.abs(rZ[ok])
    df = np.abs(fug[ok] - rfug[ok]) / np.maximum(np.abs(rfug[ok]), 1e-300)
    print(f"CPA EOS: {int(ok.sum())} states, max rel Z {dz.max():.2e}, fugacity {df.max():.2e}")
    assert dz.max() <= 1e-10 and df.max() <= 1e-10


def _cpa_without_association():
    """H2O/CO2/nC10 PR-CPA with every association site removed and the cubic
    parameters of PR78: a0 = 0.45724 R^2 Tc^2/Pc, b = 0.0778 R Tc/Pc, c1 = m(omega)."""
    d = fluids.cpa_h2o_co2_nc10()
    R = 8.31446261815324
    for i in range(3):
        tc, pc, om = d.tc[i], d.pc[i], d.acentric[i]
        d.cpa_a0[i] = 0.45724 * R * R * tc * tc / pc
        d.cpa_b[i] = 0.07780 * R * tc / pc
        d.cpa_c1[i] = (0.379642 + 1.48503 * om - 0.164423 * om * om + 0.016666 * om ** 3
                       if om > 0.49 else 0.37464 + 1.54226 * om - 0.26992 * om * om)
        d.cpa_donors[i] = d.cpa_acceptors[i] = 0
        d.cpa_epsilon[i] = d.cpa_beta[i] = 0.0
    for k in range(NN_MAX_NC * NN_MAX_NC):
        d.cpa_cross_epsilon[k] = d.cpa_cross_beta[k] = 0.0
    return d


@pytest.mark.gpu
def test_gpu_cpa_dry_limit_recovers_pr():
    """Non-associating PR-CPA is the PR78 cubic (thermo_empirical_parameters_test.cpp:698-739,
    tolerances 2e-8 on Z and 3e-8 on phi), checked against the C oracle's PR."""
    import oracle_c
    cpa = _cpa_without_association()
    pr = fluids.make_desc(["H2O", "CO2", "nC10"],
                          [[0.0, 0.10566, 0.0], [0.10566, 0.0, 0.1141], [0.0, 0.1141, 0.0]],
                          model="PR", eos_type=5)
    P, T, x, role = _cpa_states(256, 32)
    role = np.where(role == 2, 0, role).astype(np.int32)   # liquid or vapour root
    Z, lnphi, _, st = (t.cpu().numpy() for t in _flash_gpu(cpa).eos_phase(P, T, x, role))
    oZ, olnphi, ost = oracle_c.eos_batch(pr, P, T, x, role)
    ok = (st == 0) & (ost == 0)
    assert ok.mean() > 0.9
    dz = np.abs(Z[ok] - oZ[ok])
    dphi = np.abs(np.exp(lnphi[ok]) - np.exp(olnphi[ok])) / np.exp(olnphi[ok])
    print(f"dry CPA vs PR: {int(ok.sum())} states, max |dZ| {dz.max():.2e}, max rel dphi {dphi.max():.2e}")
    assert dz.max() <= 2e-8 and dphi.max() <= 3e-8


@pytest.mark.gpu
def test_gpu_cpa_single_donor_solver_matches_analytic_water():
    """With CO2's acceptor site present but zero cross-association strength, the
    single-donor Newton solver must reproduce the analytic 4C-water site fraction
    X = 2 / (1 + sqrt(1 + 4A)) (cubic_eos.hpp:1606-1623 vs :1624-1730)."""
    analytic = fluids.cpa_h2o_co2_nc10()
    analytic.cpa_acceptors[1] = 0
    analytic.cpa_epsilon[1] = analytic.cpa_beta[1] = 0.0
    analytic.cpa_cross_epsilon[1] = analytic.cpa_cross_beta[1] = 0.0
    newton = fluids.cpa_h2o_co2_nc10()
    newton.cpa_epsilon[1] = newton.cpa_beta[1] = 0.0
    newton.cpa_cross_epsilon[1] = newton.cpa_cross_beta[1] = 0.0
    P, T, x, role = _cpa_states(256, 33)
    Za, la, _, sa = (t.cpu().numpy() for t in _flash_gpu(analytic).eos_phase(P, T, x, role))
    Zn, ln_, _, sn = (t.cpu().numpy() for t in _flash_gpu(newton).eos_phase(P, T, x, role))
    assert np.array_equal(sa, sn)
    ok = sa == 0
    dz = np.abs(Za[ok] - Zn[ok]) / np.abs(Za[ok])
    dl = np.abs(la[ok] - ln_[ok])
    print(f"analytic vs single-donor: max rel dZ {dz.max():.2e}, max |dlnphi| {dl.max():.2e}")
    assert dz.max() <= 1e-9 and dl.max() <= 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["flash_nc3_pr", "flash_nc3_cpa", "flash_nc3_sw"])
def test_gpu_cooperative_pass_matches_golden(name, monkeypatch):
    """Every state through the wavefront-per-state pass (NN_COOP_ALL=1): lane-parallel
    RR3 guesses, allocation seeds and CPA density-root brackets / bisection trees
    must reproduce the reference's outputs as the lane-per-state pass does."""
    monkeypatch.setenv("NN_COOP_ALL", "1")
    g = load_case(name)
    fl = _flash_gpu(CASES[name]())
    mine = flashcmp.to_numpy(fl.flash(g["P"], g["T"], g["z"]))
    n_status, n_label, worst, wi, sel = flashcmp.compare_flash(mine, g)
    unexplained, conditioned = libm_conditioning.split_value_failures(
        flashcmp.per_state_value_diff(mine, g), sel, flashcmp.REL_TOL, name,
        gibbs=_gibbs(CASES[name](), g["P"], g["T"], mine, g), beta=mine["beta"])
    print(f"{name} (cooperative pass): status mismatches {n_status}, label mismatches {n_label}, "
          f"max rel diff {worst:.3e}; above tolerance: {len(conditioned)} libm-conditioned, "
          f"{len(unexplained)} other {unexplained}")
    assert n_status == 0 and n_label == 0
    assert not unexplained


# Published figure of the reference's CPA micro-benchmark: water/methanol sCPA
# (general multi-donor site solver), liquid root at 389.375 K, 0.233183 MPa,
# x = [0.185, 0.815]: Z = 0.002920191895 (case/five_component_eos_tuned_compare/
# RUN_REPORT.md:104, state of test/src/benchmark/cpa_phase_benchmark.cpp:144-147,
# printed to 12 decimals).
CPA_BENCH_STATE = (233183.46666657156, 389.375, (0.185, 0.815))
CPA_BENCH_Z = 0.002920191895


@pytest.mark.skipif(not refharness.available(), reason="reference build (oracle/_ref) not present")
def test_reference_build_reproduces_published_cpa_compressibility():
    P, T, x = CPA_BENCH_STATE
    ref = refharness.Reference(fluids.cpa_water_methanol(), FlashOptions.defaults())
    Z, _, _, st = ref.eos_phase(np.array([P]), np.array([T]), np.array([x]), np.array([0], np.int32))
    assert st[0] == 0 and abs(Z[0] - CPA_BENCH_Z) <= 5e-13


@pytest.mark.gpu
def test_gpu_cpa_general_site_solver_matches_published_and_reference():
    """E6 general multi-donor association (water 4C + methanol 2B): the published
    benchmark compressibility, and 512 seeded states against the reference's
    outputs (golden fixture written by tests/golden/make_golden.py)."""
    fluid = fluids.cpa_water_methanol()
    fl = _flash_gpu(fluid)
    P, T, x = CPA_BENCH_STATE
    Z, _, _, st = (t.cpu().numpy() for t in fl.eos_phase([P], [T], [x], [0]))
    assert st[0] == 0 and abs(Z[0] - CPA_BENCH_Z) <= 5e-13, Z[0]
    g = dict(np.load(GOLDEN / "eos_cpa_water_methanol.npz"))
    Z, _, fug, st = (t.cpu().numpy() for t in fl.eos_phase(g["P"], g["T"], g["x"], g["role"]))
    assert np.array_equal(st, g["status"])
    ok = g["status"] == 0
    dz = np.abs(Z[ok] - g["Z"][ok]) / np.abs(g["Z"][ok])
    df = np.abs(fug[ok] - g["fug"][ok]) / np.maximum(np.abs(g["fug"][ok]), 1e-300)
    print(f"general CPA: {int(ok.sum())} states, max rel Z {dz.max():.2e}, fugacity {df.max():.2e}")
    assert ok.sum() > len(ok) // 2 and dz.max() <= 1e-10 and df.max() <= 1e-10


@pytest.mark.skipif(not refharness.available(), reason="reference build (oracle/_ref) not present")
def test_cpa_general_fixture_matches_reference_build():
    g = dict(np.load(GOLDEN / "eos_cpa_water_methanol.npz"))
    ref = refharness.Reference(fluids.cpa_water_methanol(), FlashOptions.defaults())
    Z, _, fug, st = ref.eos_phase(g["P"], g["T"], g["x"], g["role"])
    assert np.array_equal(st, g["status"]) and np.array_equal(Z, g["Z"])


def test_host_build_of_cpa_matches_general_fixture(tmp_path):
    """The CPA code of this library compiled for the host (tests/cpa_host_check.cpp)
    reproduces the reference's general-site-solver fixture; a CPU check of the
    same source the GPU runs."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    exe = tmp_path / "cpa_host_check"
    subprocess.run([hipcc, "-std=c++17", "-O1", "-ffp-contract=off", "-x", "hip",
                    "--cuda-host-only", str(Path(__file__).parent / "cpa_host_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True)
    g = dict(np.load(GOLDEN / "eos_cpa_water_methanol.npz"))
    fd = tmp_path / "fluid.bin"
    fd.write_bytes(bytes(fluids.cpa_water_methanol()))
    st = tmp_path / "states.bin"
    with open(st, "wb") as f:
        np.array([len(g["P"])], np.int64).tofile(f)
        for k in ("P", "T", "x"):
            np.ascontiguousarray(g[k], np.float64).tofile(f)
        np.ascontiguousarray(g["role"], np.int32).tofile(f)
    out = subprocess.run([str(exe), str(fd), str(st)], capture_output=True, text=True, check=True)
    rows = np.array([[float(v) for v in line.split()] for line in out.stdout.splitlines()])
    assert np.array_equal(rows[:, 0].astype(np.int32), g["status"])
    dz = np.abs(rows[:, 1] - g["Z"]) / np.abs(g["Z"])
    assert dz.max() <= 1e-12, dz.max()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["flash_nc3_pr", "flash_nc3_sw", "flash_nc3_cpa", "flash_nc6_pr"])
def test_gpu_plain_form_is_bit_identical_to_scheduled(name):
    """The plain baseline form (one work-item per state in index order, no
    persistent queue, no sort) gives the scheduled form's results bit for bit."""
    import torch

    fluid = CASES[name]()
    P, T, z = batched_flash_states(4096, fluid, seed=13, near_critical_fraction=0.1)
    fl = _flash_gpu(fluid)
    sched = flashcmp.to_numpy(fl.flash(P, T, z))
    fl.set_schedule(plain=1)
    plain = flashcmp.to_numpy(fl.flash(P, T, z))
    torch.cuda.synchronize()
    for key in ("presence", "status", "beta", "x", "Z"):
        assert np.array_equal(np.asarray(plain[key]).view(np.uint8),
                              np.asarray(sched[key]).view(np.uint8)), key


@pytest.mark.gpu
@pytest.mark.skipif(not refharness.available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("viscosity", ["mcbride_wright", "iapws2008"])
def test_gpu_flash_with_aqueous_closures_matches_reference_build(viscosity):
    """H2O/CO2/nC10 PR with the Water-role closures of the reference's benchmark case
    (IAPWS-IF97 + Garcia molar volume; the IAPWS-2008 variant also restricts the
    Water role to <= 1e-4 solute): labels, statuses, saturations and molar
    densities against the reference build."""
    fluid = fluids.with_aqueous_closures(fluids.h2o_co2_nc10("PR"), viscosity=viscosity)
    P, T, z = batched_flash_states(1024, fluid, seed=7, near_critical_fraction=0.1)
    ref = refharness.Reference(fluid, FlashOptions.defaults()).flash(P, T, z, threads=8)
    mine = flashcmp.to_numpy(_flash_gpu(fluid).flash(P, T, z))
    n_status, n_label, worst, wi, sel = flashcmp.compare_flash(mine, ref)
    water = (ref["presence"] & 4) != 0
    print(f"aqueous closures ({viscosity}): status mismatches {n_status}, label mismatches "
          f"{n_label}, max rel diff {worst:.3e} (state {wi}); {int(water.sum())} states with water")
    assert n_status == 0 and n_label == 0
    assert worst <= flashcmp.REL_TOL
    # the closure is in force: the water phase's molar density is not P / (R T Z)
    cubic = P[water] / (8.31446261815324 * T[water] * ref["Z"][water, 2])
    assert np.min(np.abs(ref["rho_m"][water, 2] / cubic - 1.0)) > 1e-3
