"""Serving loop: each result is the same alone and in the mix.

Requests of different sizes, parameters and step counts arrive while others
are being stepped; the server admits them into its running batch.  In the
deterministic mode every request's final particle state must equal, bit for
bit, the state of the same request solved alone; the served summary must
account for every request.
"""
import numpy as np
import pytest

import mpmw
import formats
import serve
import workloads

pytestmark = pytest.mark.gpu


def small_requests():
    cfgs = sorted(workloads.batch_config2(), key=lambda c: np.prod(c["cells"]))[:4]
    steps = [40, 65, 30, 52]
    arrivals = [0.0, 0.0, 0.05, 0.1]
    return [serve.Request(i, a, dict(c, steps=s)) for i, (c, s, a) in enumerate(zip(cfgs, steps, arrivals))]


def test_served_results_equal_alone_bit_exact():
    reqs = serve.serve_batched(small_requests(), "fp32", deterministic=True, keep_results=True)
    alone = serve.serve_serial(small_requests(), "fp32", deterministic=True, keep_results=True)
    overlapped = sum(1 for q in reqs for r in reqs
                     if q is not r and q.admitted < r.finished and r.admitted < q.finished)
    assert overlapped > 0, "requests must share the device in this test"
    for q, r in zip(reqs, alone):
        assert q.done_steps == q.steps == r.steps
        for k, v in r.particles.items():
            if v is None:
                continue
            assert np.array_equal(np.asarray(q.particles[k]).view(np.uint8), np.asarray(v).view(np.uint8)), (
                q.cfg["name"], k)
        # the result message: the same NDJSON line served and alone, and the
        # restatement's line of the final state
        assert q.payload == r.payload and q.payload.startswith(b'{"count":')
        p = q.particles
        assert q.payload == formats.ndjson_particles(q.steps, p["pos"], p["temperature"], p["status"])


def test_served_summary_and_budget():
    reqs = small_requests()
    # a budget below any two requests: one resident at a time
    tight = min(serve.state_bytes(q, "fp32") for q in reqs)
    out = serve.summary(serve.serve_batched(reqs, "fp32", budget=tight))
    assert out["requests"] == len(reqs)
    assert all(q.finished >= q.admitted >= q.arrival for q in reqs)
    for q in reqs:
        for r in reqs:
            if q is not r:
                assert q.finished <= r.admitted or r.finished <= q.admitted, (q.rid, r.rid)
    assert out["p99_time_to_result_s"] >= out["p50_time_to_result_s"] > 0


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("kind", ["wedge", "voxel", "apic", "hardening"])
def test_admission_bound_covers_device_bytes(kind, precision):
    """The admission estimate is an upper bound of what a resident request
    really holds on the device (pool size classes, permutation buffers, the
    binning, tile windows, a voxel SDF, APIC and backstress fields included),
    after enough steps to rebuild the grid and re-bin."""
    cfg = sorted(workloads.batch_config2(), key=lambda c: np.prod(c["cells"]))[5]
    if kind == "voxel":
        cfg = workloads.voxel_tool_case(cfg)
    elif kind == "apic":
        cfg = dict(cfg, transfer="apic")
    elif kind == "hardening":
        cfg = dict(cfg, params=dict(cfg["params"], kinematic_hardening=2e9))
    q = serve.Request(0, 0.0, dict(cfg, steps=30))
    s = serve._new_solver(q, precision, False)
    try:
        s.run(30, cfg["dt"])
        s.synchronize()
        held = s.device_bytes()
        assert 0 < held <= serve.state_bytes(q, precision), (kind, held, serve.state_bytes(q, precision))
    finally:
        s.close()
