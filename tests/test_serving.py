"""Serving (continuous batching of member Newton steps, merged phase updates of
members that share a fluid): each step's result is the same alone and in the
mix, memory admission holds, and the server's steps match the library's direct
calls bit for bit."""
from __future__ import annotations

import numpy as np
import pytest


def _members():
    from nonenone_amd.serving import Member
    spec = [("nc3-PR", 40, 30), ("nc6-SW", 30, 20), ("nc3-CPA", 12, 10), ("nc3-SW", 50, 40),
            ("nc6-PR", 25, 25), ("nc3-PR", 20, 12)]
    return [Member(kind, nx, ny, seed=100 + i) for i, (kind, nx, ny) in enumerate(spec)]


@pytest.mark.gpu
def test_gpu_results_are_the_same_alone_and_in_the_mix():
    from nonenone_amd.serving import Server, poisson_arrivals, run_trace
    members = _members()
    # three steps of every member, all waiting at once: whole batches, and the two
    # nc3-PR members' phase updates merge; then a Poisson trace of mixed arrivals
    arrivals = [(0.0, m) for _ in range(3) for m in range(len(members))]
    arrivals += [(t + 0.05, m) for t, m in poisson_arrivals(12, len(members), 2000.0, seed=3)]
    alone = run_trace(Server(1, 8 << 30), members, arrivals, hold_start=True)
    mixed = run_trace(Server(16, 8 << 30), members, arrivals, hold_start=True)
    assert alone.stats["peak_batch_steps"] == 1 and alone.stats["peak_merged_cells"] == 0
    assert mixed.stats["peak_batch_steps"] > 1, mixed.stats
    assert mixed.stats["peak_merged_cells"] > 0, mixed.stats   # two nc3-PR members merged
    assert dict(alone.checksums) == dict(mixed.checksums)
    assert alone.nonok == mixed.nonok


@pytest.mark.gpu
def test_gpu_memory_budget_limits_admission():
    from nonenone_amd.serving import Server, run_trace
    all_members = _members()
    members = [all_members[i] for i in (0, 1, 3)]   # outputs within a factor of two
    sizes = [m.output_bytes() for m in members]
    budget = (max(sizes) + 2 * min(sizes)) // 2
    # room for one step's outputs at a time (any two need more): admission must
    # serialise the steps
    assert max(sizes) + 4096 < budget < 2 * min(sizes)
    arrivals = [(0.0, i) for i in range(len(members))]
    tight = run_trace(Server(4, budget), members, arrivals)
    assert tight.stats["peak_inflight"] == 1, tight.stats
    assert tight.stats["peak_memory_in_use"] <= budget
    roomy = run_trace(Server(4, 8 << 30), members, arrivals)
    assert roomy.stats["peak_inflight"] > 1
    assert dict(tight.checksums) == dict(roomy.checksums)


@pytest.mark.gpu
def test_gpu_server_step_matches_direct_calls():
    import torch
    from nonenone_amd.serving import Server, checksum, run_trace
    members = _members()[:2]
    trace = run_trace(Server(2, 8 << 30), members, [(0.0, 0), (0.0, 1)])
    for i, m in enumerate(members):
        m.reset()
        st, _ = m.fl.update_phase_state(m.T, m.prim, m.ps, has_well_unknown=m.has_well)
        out = m.asm.evaluate(86400.0, m.prim, m.ps, m.acc)
        torch.cuda.synchronize()
        got = dict(trace.checksums)[(i, 0)]
        assert got[0] == checksum(out["residual"])
        assert got[1] == checksum(st)
        assert got[3] == checksum(out["bsr"])


@pytest.mark.gpu
@pytest.mark.skip(reason="five- and eight-component device builds switched off after an "
                         "unexplained GPU memory fault in the five-component phase update")
def test_gpu_component_count_mix_is_the_same_alone_and_in_the_mix():
    """The serving mix of component counts 3, 5, 6 and 8 (PR, SW and PR-CPA
    members, two members per merged fluid group): every step's residual,
    Jacobian and statuses are identical whether it runs alone or batched."""
    from nonenone_amd.serving import Member, Server, run_trace
    spec = [("nc8-PR", 20, 16), ("nc5-PR", 24, 20), ("nc3-CPA", 16, 12), ("nc8-SW", 16, 12),
            ("nc5-SW", 20, 14), ("nc6-PR", 18, 18), ("nc3-SW", 30, 20), ("nc8-PR", 12, 10),
            ("nc5-PR", 14, 12)]
    members = [Member(kind, nx, ny, seed=300 + i) for i, (kind, nx, ny) in enumerate(spec)]
    arrivals = [(0.0, m) for _ in range(2) for m in range(len(members))]
    alone = run_trace(Server(1, 8 << 30), members, arrivals, hold_start=True)
    mixed = run_trace(Server(16, 8 << 30), members, arrivals, hold_start=True)
    assert mixed.stats["peak_batch_steps"] > 1, mixed.stats
    assert mixed.stats["peak_merged_cells"] > 0, mixed.stats
    assert dict(alone.checksums) == dict(mixed.checksums)
    assert alone.nonok == mixed.nonok
