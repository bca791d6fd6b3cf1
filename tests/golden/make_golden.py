"""Writes the golden flash fixtures of tests/golden/ from the reference build.

Run in a checkout that has /root/reference (oracle/_ref/libnn_ref.so built by
`make -C oracle`).  Each fixture holds seeded inputs (batched-flash workload
generator, 10% near-critical states) and the reference's own outputs of
CubicThreePhaseFlash::flash for them:

    python tests/golden/make_golden.py [--all]
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from nonenone_amd import fluids  # noqa: E402
from nonenone_amd._abi import FlashOptions  # noqa: E402
from nonenone_amd.workloads import batched_flash_states  # noqa: E402
import refharness  # noqa: E402

CASES = {
    "flash_nc3_pr": (lambda: fluids.h2o_co2_nc10("PR"), 101),
    "flash_nc3_sw": (lambda: fluids.h2o_co2_nc10("SW"), 102),
    "flash_nc6_pr": (lambda: fluids.six_component("PR"), 103),
    "flash_nc6_sw": (lambda: fluids.six_component("SW"), 104),
    "flash_nc3_cpa": (fluids.cpa_h2o_co2_nc10, 105),
}
N = 192


def assembly_fixture():
    """12x8x1 grid, H2O/CO2/nC10 PR: history flash, Newton perturbation, phase
    update, accumulation, residual and BSR Jacobian — all from the reference."""
    from nonenone_amd.assembly import assign_flash_result, perturb_newton_state
    from nonenone_amd.workloads import grid_workload

    fluid = fluids.h2o_co2_nc10("PR")
    ref = refharness.Reference(fluid, FlashOptions.defaults())
    w = grid_workload(12, 8, 1, 10.0, 10.0, 5.0, fluid, seed=5)
    fl = ref.flash(w["P"], w["T"], w["z"], threads=8)
    assert fl["converged"].all()
    hist_p, hist_ps = assign_flash_result(fl, w["P"])
    cur = perturb_newton_state(hist_p, hist_ps, 3, seed=6, push_fraction=0.1)
    upd_p, upd_ps, st, flags = ref.phase_update(w["T"], cur, hist_ps)
    assert (st == 0).all()
    n = len(w["T"])
    _, _, _, acc_prev = ref.assembly(w["topo"], w["volume"], w["porosity"], w["T"], hist_p, hist_ps, np.zeros((n, 3)), 86400.0,
                                     want_residual=False, want_bsr=False)
    res, bsr, dec, acc_cur = ref.assembly(w["topo"], w["volume"], w["porosity"], w["T"], upd_p, upd_ps, acc_prev, 86400.0)
    arrays = dict(T=w["T"], porosity=w["porosity"], volume=w["volume"], hist_primary=hist_p,
                  hist_phase_state=hist_ps, cur_primary=cur, cur_phase_state=hist_ps,
                  update_status=st, update_flags=flags, upd_primary=upd_p, upd_phase_state=upd_ps,
                  accum_prev=acc_prev, accum_cur=acc_cur, residual=res, bsr=bsr, decisions=dec)
    arrays.update({f"topo_{k}": v for k, v in w["topo"].items()})
    np.savez_compressed(HERE / "assembly_nc3_pr_12x8.npz", **arrays)
    print("assembly fixture: flags", np.bincount(flags), "presence",
          np.bincount(upd_ps[:, 0].astype(int), minlength=8))


def cpa_general_fixture():
    """Water/methanol sCPA (general multi-donor site solver): single-phase EOS
    results of the reference for seeded (P, T, x, role)."""
    fluid = fluids.cpa_water_methanol()
    ref = refharness.Reference(fluid, FlashOptions.defaults())
    rng = np.random.Generator(np.random.PCG64(17))
    n = 512
    P = rng.uniform(1e5, 5e6, n)
    T = rng.uniform(300.0, 450.0, n)
    x0 = rng.uniform(0.02, 0.98, n)
    x = np.stack([x0, 1.0 - x0], axis=1)
    role = rng.integers(0, 2, n).astype(np.int32)
    Z, lnphi, fug, st = ref.eos_phase(P, T, x, role)
    np.savez_compressed(HERE / "eos_cpa_water_methanol.npz", P=P, T=T, x=x, role=role, Z=Z,
                        lnphi=lnphi, fug=fug, status=st)


def main():
    """Writes the fixtures that are missing (all of them with --all)."""
    if not refharness.available():
        raise SystemExit("oracle/_ref/libnn_ref.so missing: run `make -C oracle` first")
    everything = "--all" in sys.argv
    if everything or not (HERE / "assembly_nc3_pr_12x8.npz").exists():
        assembly_fixture()
    if everything or not (HERE / "eos_cpa_water_methanol.npz").exists():
        cpa_general_fixture()
    for name, (make, seed) in CASES.items():
        if not everything and (HERE / f"{name}.npz").exists():
            continue
        fluid = make()
        ref = refharness.Reference(fluid, FlashOptions.defaults())
        P, T, z = batched_flash_states(N, fluid, seed=seed, near_critical_fraction=0.1)
        out = ref.flash(P, T, z, threads=8)
        np.savez_compressed(HERE / f"{name}.npz", P=P, T=T, z=z, **out)
        print(name, "converged", int(out["converged"].sum()), "/", N)


if __name__ == "__main__":
    main()
