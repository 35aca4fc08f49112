"""The launch-configuration autotuner (mpmw_set_autotune, mpmw_batch_set_autotune).

Every configuration it can choose (block order of the tile pass, the grid
update and the post pass; block size of the grid update) must compute the
same results bit for bit: checked in the deterministic mode on a batch of
three workpieces of the batched distribution over 30 steps (one grid rebuild
at step 25), and on a single solver, against the default configuration; the
tuner itself runs to completion on a batch and a solver, adopts its fastest
candidates and leaves the results unchanged.
"""
import itertools

import numpy as np
import pytest

import mpmw
import parity
import test_gpu_batch as tb

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "temperature", "plastic_strain", "stress", "damage", "strain_rate", "residual_stress",
          "status", "id")


def run_batch(cfgs, nsteps, setup):
    solvers = [parity.build_gpu(c, "fp32", True) for c in cfgs]
    b = mpmw.Batch()
    for s in solvers:
        b.add(s)
    setup(b)
    for _ in range(nsteps):
        b.step([c["dt"] for c in cfgs])
    b.synchronize()
    tuning = b.tuning()
    out = [s.get_particles() for s in solvers]
    forces = [s.metrics().contact_force for s in solvers]
    for s in list(b.members):
        b.remove(s)
    b.close()
    for s in solvers:
        s.close()
    return out, forces, tuning


def same(a, b):
    for pa, pb in zip(a, b):
        for f in FIELDS:
            assert np.array_equal(pa[f], pb[f]), f


CONFIGS = [(t, g, bs, p) for t, (g, bs), p in itertools.product((0, 1), ((1, 256), (0, 64), (1, 128)), (0, 1))]


def test_every_launch_config_is_bit_identical_on_a_batch():
    cfgs = tb.smallest(3, 30)
    ref, fref, tun = run_batch(cfgs, 30, lambda b: None)
    assert tun["state"] == 0 and (tun["tile_order"], tun["grid_order"], tun["grid_block"], tun["post_order"]) == \
        (0, 1, 256, 1)
    for t, g, bs, p in CONFIGS:
        got, f, tun = run_batch(cfgs, 30, lambda b: b.set_launch_config(mpmw.launch_config(t, g, bs, p)))
        assert (tun["tile_order"], tun["grid_order"], tun["grid_block"], tun["post_order"]) == (t, g, bs, p)
        same(ref, got)
        assert f == fref, (t, g, bs, p)


def test_batch_autotune_completes_and_keeps_results():
    """Two trials per candidate, ten candidates over three phases: tuning
    ends within the 60 steps (rebuild steps give no sample); the adopted
    configuration is the fastest median of each phase."""
    cfgs = tb.smallest(3, 60)
    ref, fref, _ = run_batch(cfgs, 60, lambda b: None)
    got, f, tun = run_batch(cfgs, 60, lambda b: b.set_autotune(2))
    print(f"\n[batch autotune] {tun}", flush=True)
    assert tun["state"] == 2
    assert all(v > 0 for v in tun["tile_ms"] + tun["grid_ms"] + tun["post_ms"])
    assert tun["tile_order"] == int(np.argmin(tun["tile_ms"]))
    gbest = [(1, 256), (1, 128), (1, 64), (0, 256), (0, 128), (0, 64)][int(np.argmin(tun["grid_ms"]))]
    assert (tun["grid_order"], tun["grid_block"]) == gbest
    assert tun["post_order"] == (1 if int(np.argmin(tun["post_ms"])) == 0 else 0)
    same(ref, got)
    assert f == fref


def test_solver_autotune_and_launch_config():
    cfg = tb.smallest(1, 60)[0]

    def run(setup):
        s = parity.build_gpu(cfg, "fp32", True)
        setup(s)
        for _ in range(60):
            s.step(cfg["dt"])
        s.synchronize()
        out, tun, f = s.get_particles(), s.tuning(), s.metrics().contact_force
        s.close()
        return out, tun, f

    ref, _, fref = run(lambda s: None)
    got, tun, f = run(lambda s: s.set_autotune(2))
    assert tun["state"] == 2
    same([ref], [got])
    assert f == fref
    got, tun, f = run(lambda s: s.set_launch_config(mpmw.launch_config(1, 0, 64, 0)))
    same([ref], [got])
    assert f == fref
    s = parity.build_gpu(cfg, "fp32", True)
    with pytest.raises(mpmw.MPMError):
        s.set_launch_config(mpmw.launch_config(0, 1, 96, 1))
    with pytest.raises(mpmw.MPMError):
        s.set_autotune(17)
    s.close()
