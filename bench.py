#!/usr/bin/env python3
"""Throughput of the per-cell phase update + Natural residual/Jacobian assembly.

One step = one Newton assembly of the whole grid, the work the reference's
NaturalPetscRuntime does per Newton iteration (runtime state update, cell cache,
formFunction + formJacobian): every cell goes through the hysteretic phase
update (stability + flash where the active set changes), AD properties, the
accumulation/fugacity/closure rows and the 7-point TPFA face fluxes, and its
block row of the BSR Jacobian is written in the cached pattern order.

Workload: 200 x 200 x 50 structured grid (2,000,000 cells, dx = dy = 10 m,
dz = 5 m), nc = 3 H2O/CO2/nC10 PR-CPA of case/h2o_co2_nc10_2d_benchmark/
case_config.hpp:106-142 (4C water, CO2 B2 cross-association, simplified
radial distribution), a well unknown per cell (11 primary variables, 11x11 BSR
blocks), hydrostatic pressure + P ~ U(8, 30) MPa, T ~ U(320, 450) K,
z ~ Dirichlet(4, 1, 1), porosity ~ U(0.1, 0.3), log-normal permeability
(median 100 mD).  History state = flash(P, T, z) written back as the primary
and phase-state vectors; the assembled state is that history plus a seeded
Newton-like perturbation (5 % of cells get an active saturation pushed below
zero, which drives the phase-removal / stability path).  Inputs are resident
in HBM before the timed region; each step restores the perturbed state with a
device copy (part of the timed work) because the phase update is in place.

    python bench.py [--gpus 1] [--steps 2] [--warmup 1] [--eos CPA|PR|SW]

Prints one JSON line.  The `cpu_baseline` leg runs the reference's own code
(the oracle's build of it, oracle/_ref) on a 50 x 50 x 50 sub-block of the
same workload with one thread per host core.

Other workloads (same JSON contract, for the figures quoted beside the headline):

    --workload flash    batched three-phase flash of 2^20 independent states
                        (half nc=3 H2O/CO2/nC10, half nc=6 H2O/CO2/C1/nC4/nC10/nC16,
                        each half split evenly between PR78 and Soreide-Whitson;
                        P ~ U(1, 40) MPa, T ~ U(300, 600) K, z ~ Dirichlet(1..1), 10 %
                        of states within 2 % of a Kay's-rule pseudo-critical point),
                        with launch shapes picked by the autotuner on an 8192-state
                        sample of each sub-batch;
                        also times the plain form (one work-item per state,
                        index order) on the same states.
    --workload serving  ensemble members (nc=3/6, PR/SW/PR-CPA, 1k-32k cells each)
                        submit 256 Newton-step requests with Poisson arrivals (200/s);
                        continuous batching (phase updates of members sharing a
                        fluid merged into one launch) vs the plain
                        one-request-at-a-time form on the same trace.
    --workload cpa_phase  the reference's published CPA micro-benchmark state
                        (water/methanol sCPA, general site solver) as 2^20 GPU
                        evaluations; vs_baseline against the published 26.58 us/phase.
    --workload learned  the learned initialiser: trains the MLP on flash results
                        of 4M nc=6 PR states of the batched-flash distribution (FP32,
                        the library's own kernels), then flashes a fresh 2^20-state
                        sample with and without its hints (BF16 inference,
                        confidence threshold calibrated for zero label changes on
                        a 65536-state validation sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "cell flash+Jacobian assemblies/s (FP64, ≤1e-10 rel); p50/p99 batch latency"
HBM_PEAK_GBS = 8000.0
DT = 86400.0
EOS_NAMES = {"CPA": "PR-CPA (4C water, CO2 B2 cross-association)", "PR": "PR78", "SW": "Soreide-Whitson"}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


FP64_PEAK_TFLOPS = 78.6
# FP64 operations of one cell's phase update on the PR-CPA headline workload
# (add, mul, transcendental = 1, fma = 2), measured by the FP64 instruction
# counters of `bench.py --workload phase_cell_ops` (256 seeded cells, one cell
# per launch sequence, so one active lane per wavefront): 1.5349e9 over 256 cells.
# profiles/phase_update_cpa_fp64_ops_single_cell_sample.md
PHASE_UPDATE_FP64_OPS_PER_CELL = 5.9956e6
# HBM bytes (L2 <-> memory, FETCH_SIZE x 2 + WRITE_SIZE x 1 against a 2 GiB copy)
# of one headline step's kernels, from the PMC passes of
# profiles/phase_update_natural_assembly_200x200x50_cpa_hbm_traffic.md (the
# assembly kernel's from its `--workload assembly_only` passes)
PHASE_UPDATE_TRAFFIC_BYTES = 2.3692e12
ASSEMBLE_TRAFFIC_BYTES = 1.3880e11


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def assemble_bytes(ncell: int, nconn: int, nnzb: int, nc: int, np_: int, nps: int) -> int:
    """Compulsory HBM bytes of one assembly launch: read primary, phase state,
    T, porosity, volume and M^n per cell and (T_ij, g dz) per connection; write
    the residual row and every BSR block once."""
    read = ncell * 8 * (np_ + nps + 3 + nc) + nconn * 16
    write = ncell * 8 * np_ + nnzb * 8 * np_ * np_
    return read + write


def build_workload(nx, ny, nz, seed=0, eos="CPA"):
    import torch

    from nonenone_amd import fluids, workloads
    from nonenone_amd.assembly import NaturalAssembly, assign_flash_result, perturb_newton_state
    from nonenone_amd.flash import ThreePhaseFlash

    fluid = fluids.cpa_h2o_co2_nc10() if eos == "CPA" else fluids.h2o_co2_nc10(eos)
    nc = fluid.nc
    t0 = time.time()
    w = workloads.grid_workload(nx, ny, nz, 10.0, 10.0, 5.0, fluid, seed=seed, hydrostatic=True)
    log(f"grid {nx}x{ny}x{nz}: {w['ncell']} cells, {int(w['topo']['conn_offsets'][-1])} "
        f"connections, {int(w['topo']['row_offsets'][-1])} blocks ({time.time() - t0:.1f}s)")
    fl = ThreePhaseFlash(fluid)
    t0 = time.time()
    keys = ("presence", "status", "beta", "saturation", "x", "Z", "k_go", "k_wo")
    parts = {k: [] for k in keys}
    chunk = 1 << 18
    for a in range(0, w["ncell"], chunk):
        b = min(a + chunk, w["ncell"])
        res = fl.flash(w["P"][a:b], w["T"][a:b], w["z"][a:b])
        for k in keys:
            parts[k].append(getattr(res, k).cpu().numpy())
        log(f"history flash {b}/{w['ncell']} ({time.time() - t0:.1f}s)")
    flash = {k: np.concatenate(v) for k, v in parts.items()}
    del parts, res
    counts = np.bincount(flash["presence"].astype(np.int64), minlength=8)
    log(f"history flash of {w['ncell']} states: {time.time() - t0:.1f}s; presence counts "
        f"{counts.tolist()}; non-OK statuses {int((flash['status'] != 0).sum())}")
    primary, ps = assign_flash_result(flash, w["P"], has_well=True)
    pert = perturb_newton_state(primary, ps, nc, seed=seed + 1)
    dev = torch.device("cuda")
    # four bottom-hole-pressure wells completed through every layer (two CO2
    # injectors, two producers), their couplings added to the cached pattern
    specs, perf_cells, perf_wi = workloads.vertical_bhp_wells(nx, ny, nz, 10.0, 10.0, 5.0,
                                                              w["perm"], w["P"], nc=nc)
    wells = workloads.make_wells(specs, nc)
    topo = workloads.pattern_add_wells(w["topo"], wells, perf_cells)
    asm = NaturalAssembly(fl, topo, w["volume"], w["porosity"], w["T"], has_well_unknown=True)
    asm.set_wells(wells, perf_cells, perf_wi, workloads.SURFACE_DENSITY)
    log(f"{len(specs)} BHP wells, {len(perf_cells)} perforations; pattern "
        f"{int(topo['row_offsets'][-1])} blocks")
    hist_p = torch.as_tensor(primary, device=dev)
    hist_s = torch.as_tensor(ps, device=dev)
    acc = asm.accumulation(hist_p, hist_s)
    del hist_p, hist_s
    state = dict(prim_pert=torch.as_tensor(pert, device=dev), ps_pert=torch.as_tensor(ps, device=dev),
                 acc=acc, T=torch.as_tensor(w["T"], device=dev))
    state["prim"] = state["prim_pert"].clone()
    state["ps"] = state["ps_pert"].clone()
    state["out"] = asm.allocate(dev)
    host = dict(pert=pert, ps=ps, w=w, acc=acc.cpu().numpy())
    return fluid, fl, asm, state, host


def run_step(fl, asm, s):
    s["prim"].copy_(s["prim_pert"])
    s["ps"].copy_(s["ps_pert"])
    st, flags = fl.update_phase_state(s["T"], s["prim"], s["ps"], has_well_unknown=True)
    asm.evaluate(DT, s["prim"], s["ps"], s["acc"], out=s["out"])
    s["status"], s["flags"] = st, flags
    return st


def cpu_baseline(fluid, host, dims, sub=(50, 50, 50), target_s=12.0, gpu=None):
    """The reference's phase update + assembly (oracle/_ref build) on a sub-block.
    `gpu` (the timed step's phase-update outputs as host arrays): also compared
    with the reference's on the sub-block's cells (parity at the bench's scale)."""
    sys.path.insert(0, str(ROOT / "tests"))
    import refharness

    from nonenone_amd import workloads
    from nonenone_amd._abi import FlashOptions

    if not refharness.available():
        log("oracle/_ref/libnn_ref.so not built: no CPU baseline")
        return None
    w = host["w"]
    nx, ny, nz = dims
    sx, sy, sz = (min(a, b) for a, b in zip(sub, dims))
    ii, jj, kk = np.meshgrid(np.arange(sx), np.arange(sy), np.arange(sz), indexing="ij")
    cells = ((kk * ny + jj) * nx + ii).transpose(2, 1, 0).reshape(-1)  # natural order of the sub-block
    topo = workloads.structured_topology(sx, sy, sz, 10.0, 10.0, 5.0, w["perm"][cells])
    ref = refharness.Reference(fluid, FlashOptions.defaults())
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    T = w["T"][cells]
    prim0 = host["pert"][cells]
    ps0 = host["ps"][cells]
    acc_prev = host["acc"][cells]
    t_total = 0.0
    passes = 0
    while t_total < target_s and passes < 50:
        t0 = time.perf_counter()
        prim, ps, _, _ = ref.phase_update(T, prim0, ps0, has_well=True, threads=cores)
        ref.assembly(topo, w["volume"][cells], w["porosity"][cells], T, prim, ps, acc_prev, DT,
                     has_well=True, threads=cores)
        t_total += time.perf_counter() - t0
        passes += 1
    n = len(cells)
    # one core: the first 4,000 cells of the sub-block (a 40 x 2 x 50 corner keeps
    # its own connections), one pass
    single = cells.reshape(sz, sy, sx)[:, :2, :40].reshape(-1)
    topo1 = workloads.structured_topology(min(40, sx), min(2, sy), sz, 10.0, 10.0, 5.0,
                                          w["perm"][single])
    t0 = time.perf_counter()
    p1, s1, _, _ = ref.phase_update(w["T"][single], host["pert"][single], host["ps"][single],
                                    has_well=True, threads=1)
    ref.assembly(topo1, w["volume"][single], w["porosity"][single], w["T"][single], p1, s1,
                 host["acc"][single], DT, has_well=True, threads=1)
    t1 = time.perf_counter() - t0
    out = {"value": n * passes / t_total, "unit": "cells/s", "cores": cores, "kind": "reference",
           "cpu_model": cpu_model(),
           "single_core_value": len(single) / t1,
           "sample": f"{sx}x{sy}x{sz} sub-block ({n} cells) of the same workload, {passes} passes "
                     f"of phase update + residual + BSR Jacobian in {t_total:.1f}s on {cores} "
                     f"threads (one per core of the box's 16-core share); single core: "
                     f"{len(single)} cells of it in {t1:.1f}s"}
    if gpu is not None:
        r_prim, r_ps, r_st, r_fl = ref.phase_update(T, prim0, ps0, has_well=True, threads=cores)
        g_prim, g_ps = gpu["prim"][cells], gpu["ps"][cells]
        d = np.abs(g_prim - r_prim) / np.maximum(np.maximum(np.abs(g_prim), np.abs(r_prim)), 1.0)
        per_cell = d.max(axis=1)
        out["parity"] = {
            "cells": int(n), "status_mismatches": int((gpu["status"][cells] != r_st).sum()),
            "flag_mismatches": int((gpu["flags"][cells] != r_fl).sum()),
            "label_mismatches": int((g_ps[:, 0] != r_ps[:, 0]).sum()),
            "primary_max_rel_diff": float(per_cell.max()),
            "cells_above_1e-10": int((per_cell > 1e-10).sum())}
    return out


FLASH_CASES = (("nc3", "PR"), ("nc3", "SW"), ("nc6", "PR"), ("nc6", "SW"))


def flash_fluid(case):
    from nonenone_amd import fluids
    comp, model = case
    return fluids.h2o_co2_nc10(model) if comp == "nc3" else fluids.six_component(model)


def flash_bytes(nc: int) -> int:
    """Compulsory HBM bytes of one flash: read P, T, z; write presence, converged,
    status, iterations, beta, S, x, Z, rho_m, K_go, K_wo."""
    return (2 + nc) * 8 + 1 + 1 + 4 + 4 + 3 * 8 * 4 + 3 * nc * 8 + 2 * nc * 8


def flash_workload(args) -> dict:
    """2^20 independent states (4 sub-batches) through nn_flash_batch."""
    import torch

    from nonenone_amd import workloads
    from nonenone_amd.flash import FlashBatchResult, ThreePhaseFlash

    total = args.states
    per = total // len(FLASH_CASES)
    # critical points of 128 seeded compositions per component count, estimated with
    # the PR flash on the GPU (workload construction, before the timed region); the
    # Soreide-Whitson sub-batches share them (SW changes only the Water role)
    loci = {}
    for comp in sorted({c for c, _ in FLASH_CASES}):
        t0 = time.time()
        pr = flash_fluid((comp, "PR"))
        fl_pr = ThreePhaseFlash(pr)
        rng = np.random.Generator(np.random.PCG64(4242))
        zs = rng.dirichlet(np.ones(pr.nc), 128)

        def host_flash(P, T, z, fl=fl_pr):
            r = fl.flash(torch.as_tensor(P, device="cuda"), torch.as_tensor(T, device="cuda"),
                         torch.as_tensor(np.ascontiguousarray(z), device="cuda"))
            return {"presence": r.presence.cpu().numpy(), "status": r.status.cpu().numpy(),
                    "x": r.x.cpu().numpy()}

        loci[comp] = workloads.critical_points(host_flash, zs, np.arange(250.0, 700.1, 5.0),
                                               np.geomspace(5.0e4, 40.0e6, 48))
        ok = loci[comp]["valid"]
        log(f"{comp}: critical points of {len(zs)} compositions in {time.time() - t0:.1f}s, "
            f"{int(ok.sum())} valid, median score "
            f"{np.median(loci[comp]['score'][ok]) if ok.any() else float('nan'):.3g}")
        del fl_pr
    subs = []
    for k, case in enumerate(FLASH_CASES):
        fluid = flash_fluid(case)
        fl = ThreePhaseFlash(fluid)
        P, T, z = workloads.batched_flash_states(per, fluid, seed=100 + k,
                                                 near_critical_fraction=0.1, locus=loci[case[0]])
        d = dict(P=torch.as_tensor(P, device="cuda"), T=torch.as_tensor(T, device="cuda"),
                 z=torch.as_tensor(z, device="cuda"))
        t0 = time.time()
        # (a Soreide-Whitson batch costs ~20x a PR one: a smaller tuning sample)
        sample = min(per, args.tune_sample // (4 if case[1] == "SW" else 1))
        best, best_ms = fl.autotune(d["P"][:sample], d["T"][:sample], d["z"][:sample], repeats=1)
        log(f"{case}: autotune on {sample} states {time.time() - t0:.1f}s -> {best} "
            f"({best_ms:.1f} ms)")
        subs.append(dict(case=case, fl=fl, d=d, host=(P, T, z), sched=best,
                         out=FlashBatchResult.empty(per, fl.nc)))

    def step(report=False):
        for sb in subs:
            ts = time.perf_counter()
            sb["fl"].flash(sb["d"]["P"], sb["d"]["T"], sb["d"]["z"], out=sb["out"])
            if report:
                torch.cuda.synchronize()
                sb["seconds"] = time.perf_counter() - ts
                log(f"  {sb['case']}: {sb['seconds']:.3f}s")

    for i in range(args.warmup):
        t0 = time.time()
        step()
        torch.cuda.synchronize()
        log(f"warmup {i}: {time.time() - t0:.3f}s")
    for sb in subs:
        sb["fl"].set_kernel_timing(True)
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        step(report=True)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
        log(f"step {i}: {time.perf_counter() - t0:.3f}s since the timed region began")
    elapsed = time.perf_counter() - t0
    kern = {}
    for sb in subs:
        ms, cnt = sb["fl"].kernel_time("flash")
        cms, _ = sb["fl"].kernel_time("flash_coop")
        kern["/".join(sb["case"])] = dict(flash_ms=ms / max(cnt, 1), coop_ms=cms / max(cnt, 1),
                                          schedule=sb["sched"])
        sb["fl"].set_kernel_timing(False)
    nonok = sum(int((sb["out"].status != 0).sum().item()) for sb in subs)
    log(f"timed {args.steps} steps in {elapsed:.3f}s; non-OK statuses {nonok}")

    # the plain form (one work-item per state, index order) on the same states (one pass)
    plain = {}
    t_plain = 0.0
    for sb in subs:
        fl = sb["fl"]
        keep = fl.schedule()
        fl.set_schedule(plain=1)
        ref_out = FlashBatchResult.empty(per, fl.nc)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        fl.flash(sb["d"]["P"], sb["d"]["T"], sb["d"]["z"], out=ref_out)
        torch.cuda.synchronize()
        dt_ = time.perf_counter() - ts
        t_plain += dt_
        fl.set_schedule(**{k: v for k, v in keep.items()})
        same = all(torch.equal(getattr(ref_out, f), getattr(sb["out"], f))
                   for f in ("presence", "status", "beta", "x", "Z"))
        plain["/".join(sb["case"])] = dict(seconds=dt_, bit_identical=same)
        log(f"plain form {sb['case']}: {dt_:.3f}s, results bit-identical: {same}")
    t_sched3 = sum(sb.get("seconds", 0.0) for sb in subs)
    per_step = elapsed / args.steps
    abytes = sum(per * flash_bytes(sb["fl"].nc) for sb in subs)
    achieved = abytes / per_step / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        cpu = flash_cpu_baseline(subs, args.cpu_sample)
        log(f"cpu baseline: {cpu}")
    return {
        "metric": "batched three-phase flashes/s (FP64)", "value": total * args.steps / elapsed,
        "unit": "flashes/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": per_step * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched flash of {total} states: nc=3 H2O/CO2/nC10 and nc=6 "
                               "H2O/CO2/C1/nC4/nC10/nC16, PR78 and Soreide-Whitson, "
                               "P~U(1,40) MPa, T~U(300,600) K, z~Dirichlet(1..1), 10% "
                               "within 2% of an estimated mixture critical point", "states": total,
                   "near_critical": {comp: {"compositions": int(len(l["z"])),
                                            "valid": int(l["valid"].sum()),
                                            "median_score": float(np.median(l["score"][l["valid"]]))
                                            if l["valid"].any() else None,
                                            "method": "upper two-phase boundary scanned and "
                                                      "refined with the PR flash; critical "
                                                      "point where the coexisting oil and gas "
                                                      "compositions coincide (workloads."
                                                      "critical_points)"}
                                     for comp, l in loci.items()}},
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "kernels_ms": kern,
        "plain_form": {"cases": "every sub-batch, one work-item per state in index order",
                       "seconds": t_plain,
                       "scheduled_seconds": t_sched3,
                       "speedup_of_scheduled": t_plain / t_sched3 if t_sched3 > 0 else None,
                       "per_case": plain},
        "roofline": {"kernel": "flashKernel (all four sub-batches)", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": abytes,
                     "note": "FP64-VALU and divergence bound; HBM bytes are the compulsory "
                             "inputs/outputs only"},
        "cpu_baseline": cpu,
    }


def flash_cpu_baseline(subs, sample: int) -> dict | None:
    """The reference's CubicThreePhaseFlash::flash (oracle/_ref) on the first `sample`
    states of every sub-batch, one thread per host core."""
    sys.path.insert(0, str(ROOT / "tests"))
    import refharness

    from nonenone_amd._abi import FlashOptions
    if not refharness.available():
        log("oracle/_ref/libnn_ref.so not built: no CPU baseline")
        return None
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    t = 0.0
    n = 0
    for sb in subs:
        ref = refharness.Reference(flash_fluid(sb["case"]), FlashOptions.defaults())
        P, T, z = (a[:sample] for a in sb["host"])
        t0 = time.perf_counter()
        ref.flash(P, T, z, threads=cores)
        t += time.perf_counter() - t0
        n += len(P)
    return {"value": n / t, "unit": "flashes/s", "cores": cores, "kind": "reference",
            "sample": f"first {sample} states of each of the 4 sub-batches ({n} flashes) "
                      f"in {t:.1f}s"}


def learned_workload(args) -> dict:
    """Train the learned initialiser on flash results, then flash a fresh sample with
    and without its hints."""
    import torch

    from nonenone_amd import fluids, learned, workloads
    from nonenone_amd.flash import ThreePhaseFlash

    fluid = fluids.six_component("PR")
    fl = ThreePhaseFlash(fluid)
    model = learned.PhaseInitModel(6, seed=7)
    t0 = time.time()
    data = learned.generate(fl, model, args.train_states, seed=10_000)
    torch.cuda.synchronize()
    t_gen = time.time() - t0
    log(f"training set: {data.X.shape[0]} converged flashes in {t_gen:.1f}s")
    t0 = time.time()
    hist = learned.train(model, data, epochs=args.epochs, batch=16384, lr=3e-3, log=log)
    torch.cuda.synchronize()
    t_train = time.time() - t0
    acc = learned.accuracy(model, data)
    log(f"trained {args.epochs} epochs in {t_train:.1f}s; training accuracy {acc:.4f}")
    Pv, Tv, zv = workloads.batched_flash_states(1 << 16, fluid, seed=77_777,
                                                near_critical_fraction=0.1)
    P, T, z = workloads.batched_flash_states(args.states, fluid, seed=99_999,
                                             near_critical_fraction=0.1)
    fl.flash(P[:4096], T[:4096], z[:4096])  # warm-up of the flash kernels
    results = {}
    cmp, _, _ = learned.evaluate(fl, model, P, T, z, modes=4, repeats=max(1, args.steps))
    cmp["threshold"] = None
    results["order"] = cmp
    log(f"hints order: {cmp}")
    for name, modes in (("order+single_phase", 5), ("order+log_k", 6)):
        th, rec = learned.calibrate(fl, model, Pv, Tv, zv, modes)
        log(f"hints {name}: calibrated threshold {th} "
            f"({ {k: v['label_mismatches'] for k, v in rec.items()} } label changes)")
        if th is None:
            results[name] = {"threshold": None}
            continue
        cmp, _, _ = learned.evaluate(fl, model, P, T, z, threshold=th, modes=modes,
                                     repeats=max(1, args.steps))
        cmp["threshold"] = th
        results[name] = cmp
        log(f"hints {name}: {cmp}")
    # the order-only hint leaves each state's arithmetic unchanged; what it changes
    # is which states share a wavefront, which the six-component device build's
    # results depend on (label_mismatches of results["order"] count it)
    best = results["order"]
    n = args.states
    return {
        "metric": "learned-order batched flashes/s (FP64, order-only hints)",
        "value": n / best["seeded_s"], "unit": "flashes/s", "n_gpus": 1,
        "steps": max(1, args.steps), "warmup": 1, "ms_per_step": best["seeded_s"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"learned initialiser: MLP 8-64-64-64-9 SiLU trained (FP32) on "
                               f"{data.X.shape[0]} flash results of nc=6 PR states, BF16 "
                               f"inference, confidence threshold calibrated on 65536 "
                               f"validation states, {n} fresh states",
                   "states": n},
        "training": {"flashes": int(data.X.shape[0]), "generate_s": t_gen, "train_s": t_train,
                     "epochs": args.epochs, "loss": hist, "accuracy": acc},
        "hints": results,
        "roofline": None, "cpu_baseline": None,
    }


CPA_PUBLISHED_US_PER_PHASE = 26.58   # RUN_REPORT.md:97-108 (median of 5 x 2000, CPU)


def cpa_phase_workload(args) -> dict:
    """The reference's published CPA micro-benchmark on the GPU: water/methanol sCPA
    liquid phase (general multi-donor site solver) at 389.375 K, 0.233183 MPa,
    x = [0.185, 0.815], as a batch of identical evaluations, plus a batch of varied
    states (x, T, P, role) for the divergent case."""
    import torch

    from nonenone_amd import fluids
    from nonenone_amd.flash import ThreePhaseFlash

    fluid = fluids.cpa_water_methanol()
    fl = ThreePhaseFlash(fluid)
    n = args.states
    dev = torch.device("cuda")
    same = dict(P=torch.full((n,), 233183.46666657156, dtype=torch.float64, device=dev),
                T=torch.full((n,), 389.375, dtype=torch.float64, device=dev),
                x=torch.tensor([[0.185, 0.815]], dtype=torch.float64, device=dev).repeat(n, 1),
                role=torch.zeros(n, dtype=torch.int32, device=dev))
    rng = np.random.Generator(np.random.PCG64(5))
    x0 = rng.uniform(0.02, 0.98, n)
    varied = dict(P=torch.as_tensor(rng.uniform(1e5, 5e6, n), device=dev),
                  T=torch.as_tensor(rng.uniform(300.0, 450.0, n), device=dev),
                  x=torch.as_tensor(np.stack([x0, 1.0 - x0], 1), device=dev),
                  role=torch.as_tensor(rng.integers(0, 2, n).astype(np.int32), device=dev))
    res = {}
    for name, d in (("benchmark_state", same), ("varied_states", varied)):
        for _ in range(max(1, args.warmup)):
            fl.eos_phase(d["P"], d["T"], d["x"], d["role"])
        torch.cuda.synchronize()
        fl.set_kernel_timing(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            Z, _, _, st = fl.eos_phase(d["P"], d["T"], d["x"], d["role"])
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        ms, cnt = fl.kernel_time("eos")
        fl.set_kernel_timing(False)
        res[name] = dict(kernel_ms=ms / max(cnt, 1), wall_ms=wall * 1e3,
                         phases_per_s=n / wall, non_ok=int((st != 0).sum().item()),
                         z0=float(Z[0].item()))
        log(f"{name}: {res[name]}")
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "tests"))
        import refharness

        from nonenone_amd._abi import FlashOptions
        if refharness.available():
            ref = refharness.Reference(fluid, FlashOptions.defaults())
            m = 2000
            args_ = (np.full(m, 233183.46666657156), np.full(m, 389.375),
                     np.tile([0.185, 0.815], (m, 1)), np.zeros(m, np.int32))
            best = float("inf")
            for _ in range(5):
                t0 = time.perf_counter()
                ref.eos_phase(*args_)
                best = min(best, time.perf_counter() - t0)
            cpu = {"value": m / best, "unit": "phases/s", "cores": 1, "kind": "reference",
                   "sample": f"{m} evaluations of the benchmark state, best of 5 "
                             f"({best / m * 1e6:.2f} us/phase; published "
                             f"{CPA_PUBLISHED_US_PER_PHASE} us/phase)"}
            log(f"cpu baseline: {cpu}")
    value = res["benchmark_state"]["phases_per_s"]
    published = 1e6 / CPA_PUBLISHED_US_PER_PHASE
    return {
        "metric": "CPA single-phase EOS evaluations/s (water/methanol sCPA, general site solver, FP64)",
        "value": value, "unit": "phases/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["benchmark_state"]["wall_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": value / published,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n} water/methanol sCPA liquid-phase evaluations at 389.375 K, "
                               "0.233183 MPa, x=[0.185,0.815] (the reference's CPA "
                               "micro-benchmark state)", "states": n},
        "published_baseline": {"us_per_phase": CPA_PUBLISHED_US_PER_PHASE,
                               "phases_per_s": published,
                               "source": "case/five_component_eos_tuned_compare/RUN_REPORT.md:97-108"},
        "batches": res,
        "roofline": None, "cpu_baseline": cpu,
    }


SERVING_KINDS = ("nc3-PR", "nc6-PR", "nc3-CPA", "nc3-SW", "nc6-SW")


def serving_workload(args) -> dict:
    """Ensemble members on one GPU: Poisson step requests, continuous batching vs the
    plain one-request-at-a-time form on the same arrival trace."""
    from nonenone_amd.serving import Member, Server, poisson_arrivals, run_trace

    rng = np.random.Generator(np.random.PCG64(2024))
    members = []
    t0 = time.time()
    for i in range(args.members):
        kind = SERVING_KINDS[i % len(SERVING_KINDS)]
        # the costlier fluids get smaller grids (PR-CPA ~30x, SW ~5x the PR cost)
        cap = 4096 if "CPA" in kind else (8192 if "SW" in kind else None)
        hi = args.max_member_cells if cap is None else min(args.max_member_cells, cap)
        cells = int(np.exp(rng.uniform(np.log(1000), np.log(hi))))
        nx = 50
        ny = max(1, (cells + nx - 1) // nx)
        t1 = time.time()
        members.append(Member(kind, nx, ny, seed=7000 + i))
        log(f"member {i} ({kind}, {nx * ny} cells) set up in {time.time() - t1:.1f}s")
    total_cells = sum(m.ncell for m in members)
    log(f"{len(members)} members, {total_cells} cells in total ({time.time() - t0:.1f}s set-up)")
    arrivals = poisson_arrivals(args.requests, len(members), rate=args.rate, seed=11)
    budget = int(args.budget_gb * (1 << 30))
    # warm-up: every member once (first launches allocate each context's queues)
    tw = time.time()
    run_trace(Server(args.batch_steps, budget), members, [(0.0, i) for i in range(len(members))])
    log(f"warm-up (every member once) {time.time() - tw:.1f}s")
    plain = run_trace(Server(1, budget), members, arrivals)
    log(f"plain form: {plain.cells} cell-steps in {plain.makespan_s:.3f}s, "
        f"p50 {np.percentile(plain.latencies_ms, 50):.1f} ms, "
        f"p99 {np.percentile(plain.latencies_ms, 99):.1f} ms")
    cont = run_trace(Server(args.batch_steps, budget), members, arrivals)
    pc, cc = dict(plain.checksums), dict(cont.checksums)
    same = pc == cc
    differing = sorted({members[k[0]].kind for k in pc if pc[k] != cc.get(k)})
    n_diff = sum(1 for k in pc if pc[k] != cc.get(k))
    if not same:
        log(f"steps whose results differ between the two forms: {n_diff} of {len(pc)} "
            f"(member kinds {differing})")
    log(f"continuous batching (<= {args.batch_steps} steps per batch): {cont.cells} cell-steps in "
        f"{cont.makespan_s:.3f}s, p50 {np.percentile(cont.latencies_ms, 50):.1f} ms, "
        f"p99 {np.percentile(cont.latencies_ms, 99):.1f} ms; stats {cont.stats}; "
        f"results identical to the plain form: {same}")
    value = cont.cells / cont.makespan_s
    plain_rate = plain.cells / plain.makespan_s
    return {
        "metric": "served cell flash+Jacobian assemblies/s (FP64); p50/p99 request latency",
        "value": value, "unit": "cells/s", "n_gpus": 1, "steps": len(arrivals), "warmup": 1,
        "ms_per_step": cont.makespan_s * 1e3 / len(arrivals), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"serving: {len(members)} members ({', '.join(SERVING_KINDS)}; "
                               f"1k-{args.max_member_cells} cells each; PR-CPA <= 4k, SW <= 8k), "
                               f"{len(arrivals)} Newton-step requests (phase update + residual "
                               f"+ BSR Jacobian), Poisson arrivals at {args.rate}/s, "
                               f"continuous batching (<= {args.batch_steps} steps per batch, "
                               f"phase updates merged per fluid), {args.budget_gb} GB budget",
                   "members": len(members), "member_cells": total_cells},
        "results_identical_to_plain_form": same,
        "steps_differing_from_plain_form": n_diff, "kinds_differing": differing,
        "p50_ms": float(np.percentile(cont.latencies_ms, 50)),
        "p99_ms": float(np.percentile(cont.latencies_ms, 99)),
        "plain_form": {"value": plain_rate, "p50_ms": float(np.percentile(plain.latencies_ms, 50)),
                       "p99_ms": float(np.percentile(plain.latencies_ms, 99)),
                       "makespan_s": plain.makespan_s},
        "speedup_vs_plain": value / plain_rate,
        "results_identical_to_plain": same,
        "server_stats": cont.stats,
        "roofline": None, "cpu_baseline": None,
    }


def phase_cell_ops_workload(args) -> dict:
    """Single-cell phase updates of a seeded sample of the assembly workload's
    cells, one cell per launch sequence: with one active lane per wavefront the
    wave-level FP64 instruction counters of a rocprofv3 --pmc run of this mode are
    the sampled cells' own FP64 operation counts (no divergence, no idle lanes).
    Same distribution as the headline step (history flash + seeded Newton
    perturbation on an nx x ny x nz grid of the same fields)."""
    import torch

    fluid, fl, asm, s, host = build_workload(args.nx, args.ny, args.nz, eos=args.eos)
    rng = np.random.Generator(np.random.PCG64(12345))
    cells = np.sort(rng.choice(asm.ncell, args.samples, replace=False))
    T, prim, ps = s["T"], s["prim_pert"], s["ps_pert"]
    counts = np.zeros(5, np.int64)
    t0 = time.time()
    for k, c in enumerate(cells):
        p1 = prim[c:c + 1].clone()
        s1 = ps[c:c + 1].clone()
        st, flags = fl.update_phase_state(T[c:c + 1], p1, s1, has_well_unknown=True)
        counts[int(st[0].item())] += 1
        if (k + 1) % 64 == 0:
            log(f"single-cell phase updates {k + 1}/{len(cells)} ({time.time() - t0:.1f}s)")
    torch.cuda.synchronize()
    return {"metric": "single-cell phase updates (FP64 operation count sample)",
            "value": len(cells) / (time.time() - t0), "unit": "cells/s", "n_gpus": 1,
            "steps": 1, "warmup": 0, "higher_is_better": True, "dtype": "f64",
            "data": "synthetic", "vs_baseline": None,
            "config": {"workload": f"{len(cells)} seeded cells of the {args.nx}x{args.ny}x"
                                   f"{args.nz} {EOS_NAMES[args.eos]} assembly workload, one "
                                   "cell per nn_phase_update_batch call", "cells": len(cells)},
            "statuses": counts.tolist(), "roofline": None, "cpu_baseline": None}


def assembly_only_workload(args) -> dict:
    """The headline step's assembly alone (cell records + residual + BSR Jacobian)
    on the phase-updated state, timed over --steps evaluations: the launches a
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE pass of the assembly kernel covers."""
    import torch

    fluid, fl, asm, s, host = build_workload(args.nx, args.ny, args.nz, eos=args.eos)
    run_step(fl, asm, s)
    torch.cuda.synchronize()
    fl.set_kernel_timing(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        asm.evaluate(DT, s["prim"], s["ps"], s["acc"], out=s["out"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms, count = fl.kernel_time("assemble")
    rms, rcount = fl.kernel_time("cell_record")
    fl.set_kernel_timing(False)
    nc = fluid.nc
    abytes = assemble_bytes(asm.ncell, asm.nconn, asm.nnzb, nc, asm.np_, asm.nps)
    avg = ms / max(count, 1) * 1e-3
    log(f"{args.steps} assemblies: assemble {ms / max(count, 1):.2f} ms, cell_record "
        f"{rms / max(rcount, 1):.2f} ms per launch")
    return {"metric": "Natural assembly evaluations/s", "value": args.steps / elapsed,
            "unit": "assemblies/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
            "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
            "dtype": "f64", "data": "synthetic", "vs_baseline": None,
            "config": {"workload": f"assembly of the {args.nx}x{args.ny}x{args.nz} "
                                   f"{EOS_NAMES[args.eos]} step", "cells": asm.ncell},
            "kernels_ms": {"assemble": ms / max(count, 1), "cell_record": rms / max(rcount, 1)},
            "roofline": {"kernel": "assembleKernel<3,true>", "bound": "hbm",
                         "achieved": abytes / avg / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": abytes / avg / 1e9 / HBM_PEAK_GBS, "traffic": None,
                         "algorithmic_bytes": abytes},
            "cpu_baseline": None}


def copy_workload(args) -> dict:
    """Device-to-device copy of 2 GiB, three times: the calibration of the HBM
    byte counters (FETCH_SIZE / WRITE_SIZE per dispatch against 2 GiB read and
    2 GiB written)."""
    import torch

    n = (2 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 3
    return {"metric": "device copy bandwidth", "value": 2 * (2 << 30) / dt / 1e9,
            "unit": "GB/s", "n_gpus": 1, "steps": 3, "warmup": 0, "higher_is_better": True,
            "dtype": "f64", "data": "synthetic", "vs_baseline": None,
            "config": {"workload": "2 GiB device-to-device copy (counter calibration)",
                       "bytes_read": 2 << 30, "bytes_written": 2 << 30},
            "roofline": None, "cpu_baseline": None}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=200)
    ap.add_argument("--ny", type=int, default=200)
    ap.add_argument("--nz", type=int, default=50)
    ap.add_argument("--eos", choices=("CPA", "PR", "SW"), default="CPA")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain-form", action="store_true",
                    help="skip the one plain-form step after the timed steps")
    ap.add_argument("--workload", choices=("assembly", "flash", "learned", "serving", "cpa_phase",
                                           "phase_cell_ops", "assembly_only", "copy"),
                    default="assembly")
    ap.add_argument("--samples", type=int, default=256)
    ap.add_argument("--tune-sample", type=int, default=4096)
    ap.add_argument("--members", type=int, default=48)
    ap.add_argument("--max-member-cells", type=int, default=32768)
    ap.add_argument("--requests", type=int, default=192)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--batch-steps", type=int, default=64)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--states", type=int, default=1 << 20)
    ap.add_argument("--cpu-sample", type=int, default=2048)
    ap.add_argument("--train-states", type=int, default=4 << 20)
    ap.add_argument("--epochs", type=int, default=6)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the path is per-cell and shards by slabs of the grid "
              "with a halo exchange that this build does not implement; run with --gpus 1.",
              file=sys.stderr)
        return 2

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 1
    if args.workload != "assembly":
        line = {"flash": flash_workload, "learned": learned_workload,
                "serving": serving_workload, "cpa_phase": cpa_phase_workload,
                "phase_cell_ops": phase_cell_ops_workload,
                "assembly_only": assembly_only_workload,
                "copy": copy_workload}[args.workload](args)
        print(json.dumps(line), flush=True)
        return 0
    fluid, fl, asm, s, host = build_workload(args.nx, args.ny, args.nz, eos=args.eos)
    ncell, nconn, nnzb = asm.ncell, asm.nconn, asm.nnzb
    for i in range(args.warmup):
        t0 = time.time()
        st = run_step(fl, asm, s)
        torch.cuda.synchronize()
        log(f"warmup {i}: {time.time() - t0:.3f}s; phase-update non-OK "
            f"{int((st != 0).sum().item())}; assembly non-OK "
            f"{int((s['out']['status'] != 0).sum().item())}")
    queue = None
    if args.warmup > 0:
        queue = fl.queue_counts()
        log(f"phase-update cells: {queue['recorded']} recorded for the lockstep trials, "
            f"{queue['reflashed']} re-flashed, {queue['deferred']} on the wavefront-per-cell pass")

    fl.set_kernel_timing(True)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record()
        run_step(fl, asm, s)
        ends[i].record()
        ends[i].synchronize()   # progress line per step (the steps run back to back)
        log(f"step {i}: {time.perf_counter() - t0:.3f}s since the timed region began")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    lat = np.array([a.elapsed_time(b) for a, b in zip(starts, ends)])
    kern = {}
    for kind in ("phase_update", "phase_update_coop", "cell_record", "assemble"):
        ms, count = fl.kernel_time(kind)
        kern[kind] = dict(total_ms=ms, launches=count, avg_ms=ms / max(count, 1))
    fl.set_kernel_timing(False)
    log(f"timed {args.steps} steps in {elapsed:.3f}s; kernels "
        + ", ".join(f"{k} {v['avg_ms']:.2f} ms" for k, v in kern.items()))

    nc = fluid.nc
    np_, nps = asm.np_, asm.nps
    abytes = assemble_bytes(ncell, nconn, nnzb, nc, np_, nps)
    avg_s = kern["assemble"]["avg_ms"] * 1e-3
    achieved = abytes / avg_s / 1e9
    roofline_assembly = {"kernel": "assembleKernel<3,true>", "bound": "hbm",
                         "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": ASSEMBLE_TRAFFIC_BYTES,
                         "algorithmic_bytes": abytes}
    # the dominant kernels: the phase update's passes (record, lockstep trials, replay,
    # re-flash; cooperative), FP64 work per cell x cells / their measured time
    pu_s = (kern["phase_update"]["avg_ms"] + kern["phase_update_coop"]["avg_ms"]) * 1e-3
    eos_cpa = args.eos == "CPA"
    flops = PHASE_UPDATE_FP64_OPS_PER_CELL * ncell if eos_cpa else None
    tflops = flops / pu_s / 1e12 if eos_cpa else None
    roofline = {"kernel": "phase update: phaseUpdateKernel + stabilityTrialKernel + "
                          "phaseUpdateReplayKernel + phaseUpdateReflashKernel + "
                          "phaseUpdateCoopKernel",
                "bound": "mfma", "achieved": tflops, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": tflops / FP64_PEAK_TFLOPS if eos_cpa else None,
                "traffic": PHASE_UPDATE_TRAFFIC_BYTES if eos_cpa else None,
                "algorithmic_flops": flops,
                "note": "FP64 vector (VALU) work; MI355X's FP64 vector and matrix peaks are both "
                        "78.6 TFLOP/s and no MFMA is used. Work per cell measured by FP64 "
                        "instruction counters on single-cell launches (PR-CPA only); traffic "
                        "is L2<->memory bytes of the passes (PMC), mostly per-lane scratch"}

    gpu = {"prim": s["prim"].cpu().numpy(), "ps": s["ps"].cpu().numpy(),
           "status": s["status"].cpu().numpy(), "flags": s["flags"].cpu().numpy()}
    plain_form = None
    if not args.no_plain_form:
        # the same step in the plain form (nn_schedule.plain: one work-item per cell
        # in index order, no condition sort, no lockstep stability trials), once
        sched_out = {k: v.cpu().numpy() for k, v in s["out"].items()}
        keep = fl.schedule()
        fl.set_schedule(plain=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_step(fl, asm, s)
        torch.cuda.synchronize()
        plain_s = time.perf_counter() - t0
        fl.set_schedule(**keep)
        same = (np.array_equal(s["status"].cpu().numpy(), gpu["status"]) and
                np.array_equal(s["flags"].cpu().numpy(), gpu["flags"]) and
                np.array_equal(s["prim"].cpu().numpy(), gpu["prim"]) and
                np.array_equal(s["ps"].cpu().numpy(), gpu["ps"]) and
                all(np.array_equal(s["out"][k].cpu().numpy(), v, equal_nan=v.dtype.kind == "f")
                    for k, v in sched_out.items()))
        plain_form = {"ms_per_step": plain_s * 1e3,
                      "speedup_of_scheduled": plain_s * 1e3 / (elapsed * 1e3 / args.steps),
                      "bit_identical": bool(same),
                      "form": "one work-item per cell in index order (ceil(n/64) blocks), no "
                              "condition sort, no lockstep stability trials; cooperative pass "
                              "for the cells past the lane budget; same assembly"}
        log(f"plain form: {plain_s:.3f}s per step, results bit-identical: {same}")

    cpu = None
    if not args.no_cpu_baseline:
        t0 = time.time()
        cpu = cpu_baseline(fluid, host, (args.nx, args.ny, args.nz), gpu=gpu)
        log(f"cpu baseline {time.time() - t0:.1f}s: {cpu}")

    ms_per_step = elapsed * 1e3 / args.steps
    line = {
        "metric": METRIC, "value": ncell * args.steps / elapsed, "unit": "cells/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"natural assembly {args.nx}x{args.ny}x{args.nz} "
                               f"({ncell} cells), nc=3 H2O/CO2/nC10 {EOS_NAMES[args.eos]}, "
                               f"7-point TPFA with gravity, hydrostatic P, well unknown per "
                               f"cell (N_p=11), 4 BHP wells (2 CO2 injectors, 2 producers) "
                               f"on {args.nz}-cell completions, phase update + residual + BSR "
                               f"Jacobian ({nnzb} 11x11 blocks)",
                   "eos": EOS_NAMES[args.eos], "cells": ncell,
                   "connections": nconn, "bsr_blocks": nnzb, "n_primary": np_},
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "kernels_ms": {k: v["avg_ms"] for k, v in kern.items()},
        "phase_update_cells": queue,
        "roofline": roofline,
        "roofline_assembly": roofline_assembly,
        "plain_form": plain_form,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
