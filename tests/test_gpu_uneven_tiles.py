"""The fused tile pass on tiles whose cells hold uneven particle counts.

A tile walks its cells slot by slot, one thread per cell; when the counts
differ, a cell's particle in slot s sits at a slot-major position found
from the counts of the cells before it.  Tiles of fewer than 32 slots take
those counts from a per-wave table built once (tile_block's s_tab), fuller
ones through a ballot and a barrier per slot.  Both are checked against the
FP64 oracle stepped from the GPU's own state (tolerance 1e-5 normwise in
FP32, 1e-9 in FP64, no status differing):
* particles displaced from the jittered lattice by up to 0.45 cell widths
  per axis (cells of uneven counts in almost every tile: the table);
* the same with the particles of 2 x 2 x 2 cells gathered into one cell
  (about 70 particles: the per-slot path), also in the deterministic mode, where
  a rerun must repeat the run bit for bit;
* a batch of a displaced and a lattice slab.
"""
import numpy as np
import pytest

import mpmw
import parity
import workloads

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "temperature", "plastic_strain", "stress", "damage", "strain_rate")
H = workloads.H


def displaced(seed=11, amount=0.45, crowd=False):
    def mutate(d):
        rng = np.random.default_rng(seed)
        p = d["pos"]
        p += rng.uniform(-amount, amount, p.shape) * H
        if crowd:
            # the particles of grid cells [g, g + 2) per axis into grid cell
            # g (the grid's origin is the particles' lower bound less 4 cells)
            org = p.min(axis=0) - 4.0 * H
            c = np.floor((p - org) / H)
            g = np.floor((np.array((10.5, 10.5, 4.5)) * H - org) / H)
            sel = np.all((c >= g) & (c < g + 2), axis=1)
            p[sel] = org + (g + rng.uniform(0.2, 0.8, (int(sel.sum()), 3))) * H
    return mutate


def grid_cell_counts(d):
    org = d["pos"].min(axis=0) - 4.0 * H
    _, counts = np.unique(np.floor((d["pos"] - org) / H).astype(int), axis=0, return_counts=True)
    return counts


def uneven_case(steps, crowd=False):
    cfg = dict(workloads.slab_config0())
    cfg["steps"] = steps
    cfg["mutate"] = displaced(crowd=crowd)
    cfg["name"] = "config0_rake_slab_displaced" + ("_crowded" if crowd else "")
    return cfg


def shadow(cfg, precision, nsteps, deterministic=False):
    """Per-step errors of the GPU step against the oracle stepped from the
    GPU's state before the step, and the GPU's final state."""
    o, d = parity.build_oracle(cfg)
    s = parity.build_gpu(cfg, precision, deterministic, particles=d)
    per_step = []
    for _ in range(nsteps):
        o.replace_particles(s.get_particles())
        o.step(cfg["dt"])
        s.step(cfg["dt"])
        per_step.append(parity.compare_particles(s.get_particles(), o.get_particles()))
    diag = s.diagnostics()
    final = s.get_particles()
    s.close()
    return per_step, diag, final


def check(per_step, tol, label):
    worst = {f: 0.0 for f in FIELDS}
    for k, (e, m) in enumerate(per_step):
        assert m == 0, (label, k)
        for f in FIELDS:
            assert e[f] <= tol, (label, k, f, e[f])
            worst[f] = max(worst[f], e[f])
    print(f"\n[{label}] worst over {len(per_step)} steps " + " ".join(f"{k}={v:.2e}" for k, v in worst.items()),
          flush=True)


def test_displaced_particles_fp32_30_steps():
    """30 steps cross the grid rebuild at step 25."""
    per_step, diag, _ = shadow(uneven_case(30), "fp32", 30)
    assert diag["max_tile_particles"] > 0
    check(per_step, 1e-5, "uneven fp32")


def test_displaced_particles_fp64():
    per_step, _, _ = shadow(uneven_case(3), "fp64", 3)
    check(per_step, 1e-9, "uneven fp64")


def test_crowded_cell_per_slot_path():
    cfg = uneven_case(4, crowd=True)
    _, d = parity.build_oracle(cfg)
    assert grid_cell_counts(d).max() >= 32
    per_step, _, _ = shadow(cfg, "fp32", 4)
    check(per_step, 1e-5, "crowded fp32")


@pytest.mark.parametrize("crowd", [False, True])
def test_deterministic_uneven_rerun_bit_exact(crowd):
    cfg = uneven_case(6, crowd=crowd)
    a, _, fa = shadow(cfg, "fp32", 6, deterministic=True)
    b, _, fb = shadow(cfg, "fp32", 6, deterministic=True)
    check(a, 1e-5, f"uneven det crowd={crowd}")
    for k in ("pos", "vel", "temperature", "plastic_strain", "stress", "damage", "strain_rate", "status", "id"):
        assert np.array_equal(np.asarray(fa[k]).view(np.uint8), np.asarray(fb[k]).view(np.uint8)), k


def test_batch_of_displaced_and_lattice_slabs():
    """Two batched steps of a displaced slab and a lattice slab against the
    oracle from the same states."""
    cfgs = [uneven_case(2), dict(workloads.slab_config0(), steps=2)]
    solvers, oracles = [], []
    for c in cfgs:
        o, d = parity.build_oracle(c)
        solvers.append(parity.build_gpu(c, "fp32", particles=d if c.get("mutate") else None))
        oracles.append(o)
    b = mpmw.Batch()
    for s in solvers:
        b.add(s)
    for step in range(2):
        for s, o in zip(solvers, oracles):
            o.replace_particles(s.get_particles())
            o.step(cfgs[0]["dt"])
        b.step([c["dt"] for c in cfgs])
        b.synchronize()
        for i, (s, o) in enumerate(zip(solvers, oracles)):
            e, m = parity.compare_particles(s.get_particles(), o.get_particles())
            assert m == 0, (step, i)
            for f in FIELDS:
                assert e[f] <= 1e-5, (step, i, f, e[f])
    for s in list(b.members):
        b.remove(s)
    b.close()
    for s in solvers:
        s.close()
