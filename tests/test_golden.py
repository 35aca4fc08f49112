"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the serial FP64 oracle reproduces every fixture bit for bit, so the
fixtures pin the oracle against regressions.  GPU: the HIP solver, in the
atomic and the deterministic mode, matches the fixtures within the parity
tolerances (1e-9 FP64; 1e-5 FP32 on the gated fields), read straight from the
files (no oracle run).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402
import parity  # noqa: E402

CASES = make_golden.cases()


def _load(name):
    with np.load(make_golden.path(name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_golden(name):
    gold = _load(name)
    out = make_golden.run_oracle(CASES[name])
    for k, v in gold.items():
        assert np.array_equal(out[k], v), f"{name}: {k} differs from the fixture"


def _compare(name, gold, s):
    p = s.get_particles()
    gi = np.argsort(p["id"], kind="stable")
    oi = np.argsort(gold["p_id"], kind="stable")
    assert np.array_equal(p["id"][gi], gold["p_id"][oi])
    same = p["status"][gi] == gold["p_status"][oi]
    errs = {}
    for f in ("pos", "vel", "temperature", "plastic_strain", "stress", "damage"):
        errs[f] = parity.normwise(p[f][gi][same], gold[f"p_{f}"][oi][same])
    g = s.grid_info()
    assert (g.nx, g.ny, g.nz) == tuple(gold["g_dims"]), name
    snap = s.grid_snapshot()
    errs["grid_mass"] = parity.normwise(snap["mass"], gold["g_mass"])
    errs["grid_momentum"] = parity.normwise(snap["mom"], gold["g_mom"])
    return errs, int((~same).sum())


@pytest.mark.gpu
@pytest.mark.parametrize("deterministic", [False, True])
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_fp64_matches_golden(name, deterministic):
    cfg = CASES[name]
    gold = _load(name)
    s = parity.build_gpu(cfg, "fp64", deterministic)
    s.set_grid_snapshot(True)
    for _ in range(cfg["steps"]):
        s.step(cfg["dt"])
    errs, mism = _compare(name, gold, s)
    print(f"\n[golden {name} fp64 det={deterministic}] mismatch={mism} "
          + " ".join(f"{k}={v:.2e}" for k, v in sorted(errs.items())), flush=True)
    assert mism == 0
    for k, v in errs.items():
        assert v <= 1e-9, (k, v)
