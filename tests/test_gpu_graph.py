"""HIP-graph replay of the step (mpmw_set_graph, mpmw_batch_set_graph).

The copy of a step's work items and its kernels are captured once per launch
configuration and replayed; the per-step values travel in the copied work
items.  In the deterministic mode a graph-stepped run must equal the
eagerly stepped run bit for bit at every step, with the program's own eager
work (downloads, synchronisations, a kernel-timed step) between the replays;
in the atomic mode it must match the oracle like the eager step.
"""
import numpy as np
import pytest

import mpmw
import parity
import workloads

pytestmark = pytest.mark.gpu


def _bits(d):
    return {k: np.ascontiguousarray(v).tobytes() for k, v in d.items() if v is not None}


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_graph_step_equals_eager_every_step(prec):
    """60 steps of the drill on a small slab (grid rebuilds at 25 and 50,
    re-binning, the chip pass every other step): the graph-stepped solver's
    particles after every step equal the eager solver's bit for bit; between
    replays the test downloads the state (synchronising), and one step in
    the middle runs with kernel timing on (eager launches)."""
    cfg = workloads.scaled(workloads.drill_config1(), (24, 16, 10), steps=60, name="graph_drill")
    a = parity.build_gpu(cfg, prec, deterministic=True)
    b = parity.build_gpu(cfg, prec, deterministic=True)
    b.set_graph(True)
    a.set_grid_snapshot(True)
    b.set_grid_snapshot(True)
    for k in range(cfg["steps"]):
        if k == 31:
            b.set_kernel_timing(True)
        if k == 32:
            b.set_kernel_timing(False)
        a.step(cfg["dt"])
        b.step(cfg["dt"])
        pa, pb = _bits(a.get_particles()), _bits(b.get_particles())
        for f in pa:
            assert pa[f] == pb[f], (k, f)
    ga, gb = _bits(a.grid_snapshot()), _bits(b.grid_snapshot())
    for f in ga:
        assert ga[f] == gb[f], f
    ma, mb = a.metrics(), b.metrics()
    for f in ("max_stress", "max_temperature", "kinetic_energy", "contact_force"):
        assert getattr(ma, f) == getattr(mb, f), f
    st = b.graph_stats()
    print(f"\n[graph {prec}] {st}", flush=True)
    assert st["replays"] >= 40 and 1 <= st["captures"] <= 12
    assert a.graph_stats()["replays"] == 0
    a.close()
    b.close()


def test_graph_step_matches_oracle_atomic():
    """Atomic mode (float atomics for particles that left their tile): the
    graph-stepped FP32 solver against the FP64 oracle at the parity bar."""
    cfg = parity.small_case()
    o, _ = parity.build_oracle(cfg)
    s = parity.build_gpu(cfg, "fp32")
    s.set_graph(True)
    s.set_grid_snapshot(True)
    for _ in range(30):
        o.step(cfg["dt"])
        s.step(cfg["dt"])
    errs, mism = parity.compare_particles(s.get_particles(), o.get_particles())
    gs, osn = s.grid_snapshot(), o.grid_snapshot()
    errs["grid_mass"] = parity.normwise(gs["mass"], osn["mass"])
    errs["grid_momentum"] = parity.normwise(gs["mom"], osn["mom"])
    print("\n[graph atomic fp32] " + " ".join(f"{k}={v:.2e}" for k, v in sorted(errs.items())), flush=True)
    assert mism == 0 and s.graph_stats()["replays"] > 0
    for k in ("pos", "temperature", "plastic_strain", "grid_mass", "grid_momentum"):
        assert errs[k] <= 1e-5, (k, errs[k])
    s.close()


def test_batch_graph_equals_eager_with_joins_and_leaves():
    """A deterministic batch stepped through graphs against the same batch
    stepped eagerly: members join at step 10 and leave at step 30 (the graph
    is captured anew each time), every member's final state equal bit for
    bit."""
    cfgs = sorted(workloads.batch_config2(count=8), key=lambda c: np.prod(c["cells"]))[:5]

    def run(graph):
        solvers = [parity.build_gpu(c, "fp32", deterministic=True) for c in cfgs]
        b = mpmw.Batch()
        b.set_graph(graph)
        for s in solvers[:3]:
            b.add(s)
        for k in range(45):
            if k == 10:
                for s in solvers[3:]:
                    b.add(s)
            if k == 30:
                b.remove(solvers[1])
            b.step(1e-8)
            if k % 7 == 0:
                b.synchronize()
        b.synchronize()
        out = [_bits(s.get_particles()) for s in solvers]
        stats = b.graph_stats()
        b.close()
        for s in solvers:
            s.close()
        return out, stats

    eager, se = run(False)
    graph, sg = run(True)
    print(f"\n[batch graph] {sg}", flush=True)
    assert se["replays"] == 0 and sg["replays"] >= 30 and sg["captures"] >= 3
    for i, (pe, pg) in enumerate(zip(eager, graph)):
        for f in pe:
            assert pe[f] == pg[f], (i, f)
