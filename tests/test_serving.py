"""Serving (continuous batching of member Newton steps): each step's result is the
same alone and in the mix, memory admission holds, and the server's steps match
the library's direct calls bit for bit."""
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
    arrivals = poisson_arrivals(18, len(members), rate=2000.0, seed=3)
    alone = run_trace(Server(1, 8 << 30), members, arrivals)
    mixed = run_trace(Server(4, 8 << 30), members, arrivals)
    assert alone.stats["peak_inflight"] == 1
    assert mixed.stats["peak_inflight"] > 1, mixed.stats
    assert dict(alone.checksums) == dict(mixed.checksums)
    assert alone.nonok == mixed.nonok


@pytest.mark.gpu
def test_gpu_memory_budget_limits_admission():
    from nonenone_amd.serving import Server, run_trace
    members = _members()[:4]
    biggest = max(m.output_bytes() for m in members)
    # room for one step's outputs at a time: admission must serialise the steps
    tight = run_trace(Server(4, int(biggest * 1.3)), members, [(0.0, i) for i in range(4)])
    assert tight.stats["peak_inflight"] == 1, tight.stats
    assert tight.stats["peak_memory_in_use"] <= int(biggest * 1.3)
    roomy = run_trace(Server(4, 8 << 30), members, [(0.0, i) for i in range(4)])
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
