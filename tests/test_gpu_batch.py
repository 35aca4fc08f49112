"""Batched stepping: a member's results are the same alone and in a batch.

Workpieces come from the batched distribution (workloads.sample_workpiece:
uneven sizes, J-C parameters within +-15 %, analytic rake wedge at different
cutting speeds and depths).  In the deterministic mode a member's particle
state after its steps must equal, bit for bit, the state of the same set-up
stepped alone, whatever joined or left the batch meanwhile.  The atomic mode
is held to the parity tolerances.  The wedge tool itself is checked against
the oracle.
"""
import numpy as np
import pytest

import mpmw
import parity
import workloads

pytestmark = pytest.mark.gpu

GATED = ("pos", "temperature", "plastic_strain")


def smallest(k, steps):
    cfgs = sorted(workloads.batch_config2(), key=lambda c: np.prod(c["cells"]))[:k]
    return [dict(c, steps=steps) for c in cfgs]


def alone(cfg, prec, det, nsteps):
    s = parity.build_gpu(cfg, prec, det)
    for _ in range(nsteps):
        s.step(cfg["dt"])
    p = s.get_particles()
    s.close()
    return p


def batched(cfgs, prec, det, nsteps, join, leave):
    """Step every set-up `nsteps` times; set-up i joins the batch at batch
    step join[i] and leaves it at batch step leave[i] (its remaining steps are
    taken alone)."""
    solvers = [parity.build_gpu(c, prec, det) for c in cfgs]
    done = [0] * len(cfgs)
    b = mpmw.Batch()
    t = 0
    while any(d < nsteps for d in done):
        for i, s in enumerate(solvers):
            if t == join[i] and done[i] < nsteps:
                b.add(s)
            if t == leave[i] and s in b.members:
                b.remove(s)
        if b.members:
            b.step([cfgs[solvers.index(s)]["dt"] for s in b.members])
            for s in b.members:
                done[solvers.index(s)] += 1
            for s in list(b.members):
                if done[solvers.index(s)] == nsteps:
                    b.remove(s)
        for i, s in enumerate(solvers):
            if s not in b.members and t >= leave[i] and done[i] < nsteps:
                s.step(cfgs[i]["dt"])
                done[i] += 1
        t += 1
    b.close()
    out = [s.get_particles() for s in solvers]
    for s in solvers:
        s.close()
    return out


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_det_batch_equals_alone_bit_exact(prec):
    cfgs = smallest(3, 60)
    got = batched(cfgs, prec, True, 60, join=[0, 0, 10], leave=[10 ** 9, 40, 10 ** 9])
    for c, g in zip(cfgs, got):
        ref = alone(c, prec, True, 60)
        for k, v in ref.items():
            if v is None:
                continue
            assert np.array_equal(np.asarray(g[k]).view(np.uint8), np.asarray(v).view(np.uint8)), (c["name"], k)


def test_atomic_batch_matches_alone():
    cfgs = smallest(3, 40)
    got = batched(cfgs, "fp64", False, 40, join=[0, 5, 0], leave=[10 ** 9, 10 ** 9, 30])
    for c, g in zip(cfgs, got):
        ref = alone(c, "fp64", False, 40)
        errs, mism = parity.compare_particles(g, ref)
        print(f"\n[batch atomic {c['name']}] mismatch={mism} "
              + " ".join(f"{k}={v:.2e}" for k, v in sorted(errs.items())), flush=True)
        assert mism == 0
        for k in GATED:
            assert errs[k] <= 1e-9, (c["name"], k, errs[k])


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_wedge_tool_matches_oracle(prec, tol):
    cfg = smallest(1, 30)[0]
    res = parity.run_case(cfg, prec)
    e = res["errors"]
    print(f"\n[wedge {cfg['name']} {prec}] cf gpu/oracle {res['metrics'][0].contact_force:.6g}/"
          f"{res['metrics'][1].contact_force:.6g} " + " ".join(f"{k}={v:.2e}" for k, v in sorted(e.items())),
          flush=True)
    assert res["status_mismatch"] == 0 and res["grid_same"]
    for k in GATED + ("grid_mass", "grid_momentum"):
        assert e[k] <= tol, (k, e[k])


def test_batch_rejects_mixed_precision_and_double_membership():
    cfg = smallest(1, 5)[0]
    a = parity.build_gpu(cfg, "fp32")
    d = parity.build_gpu(cfg, "fp64")
    b = mpmw.Batch()
    b.add(a)
    with pytest.raises(mpmw.MPMError):
        b.add(d)
    with pytest.raises(mpmw.MPMError):
        b.add(a)
    b2 = mpmw.Batch()
    with pytest.raises(mpmw.MPMError):
        b2.add(a)
    b.step(cfg["dt"])
    assert b.count() == 1
    a.close()              # destroying a member removes it from its batch
    b.step(cfg["dt"])      # empty batch: nothing to do
    b.close()
    b2.close()
    d.close()


def test_full_64_member_batch_500_steps_equals_alone():
    """The whole batched workload (64 workpieces, 2.4e7 particles) for 500
    steps in the deterministic mode: every member's end state equals, bit
    for bit, the same workpiece stepped alone for 500 steps (20 grid
    rebuilds, re-binning on cell changes)."""
    cfgs = [dict(c, steps=500) for c in workloads.batch_config2()]
    assert len(cfgs) == 64
    solvers = [parity.build_gpu(c, "fp32", True) for c in cfgs]
    b = mpmw.Batch()
    for s in solvers:
        b.add(s)
    dts = [c["dt"] for c in cfgs]
    for _ in range(500):
        b.step(dts)
    b.synchronize()
    for s in solvers:
        b.remove(s)
    b.close()
    keys = ("pos", "vel", "temperature", "plastic_strain", "stress", "damage", "status", "id")
    for i, (c, s) in enumerate(zip(cfgs, solvers)):
        got = s.get_particles()
        s.close()
        ref = alone(c, "fp32", True, 500)
        for k in keys:
            assert np.array_equal(np.asarray(got[k]).view(np.uint8), np.asarray(ref[k]).view(np.uint8)), (i, k)
    print("\n[full batch] 64 members x 500 steps equal alone, bit for bit", flush=True)
