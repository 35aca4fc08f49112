"""Write the golden fixtures: oracle (serial FP64) end states of small seeded
workloads.

The reference ships no test vectors, so these fixtures are produced by the
CPU restatement of MPMWorkpieceSolver::step from seeded inputs and pinned:
tests/test_golden.py checks that the oracle still reproduces them bit for bit
(CPU) and that the HIP solver matches them within the parity tolerances
(GPU).  Regenerate only after a deliberate change of the oracle:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = os.path.dirname(HERE)
sys.path.insert(0, TESTS)

import parity  # noqa: E402
import workloads  # noqa: E402
import oracle  # noqa: E402

PARTICLE_FIELDS = ("pos", "vel", "temperature", "plastic_strain", "stress", "damage", "strain_rate",
                   "residual_stress", "status", "id")


def translating_slab(cells_per_step, steps):
    """8x8x4-cell slab, no tool, translating at `cells_per_step` cells per
    step along x and half that along -y; tip far below (no chip pass)."""
    cfg = dict(workloads.scaled(workloads.slab_config0(), (8, 8, 4), steps=steps,
                                name=f"translating_slab_{cells_per_step}"))
    v = cells_per_step * workloads.H / cfg["dt"]
    cfg["vel"] = (v, -0.5 * v, 0.0)
    cfg["tool"] = ("none",)
    cfg["tip"] = (0.0, 0.0, -1.0)
    return cfg


def drill_small():
    """The helical drill on an 8x8x6-cell slab (2 x 2 x 1.5 mm) with a 0.5 mm
    fixture, so that the drill point engages free particles at once."""
    cfg = dict(workloads.scaled(workloads.drill_config1(), (8, 8, 6), steps=20, name="drill_8x8x6"))
    cfg["params"] = workloads.params_dict(fixture=0.5e-3)
    return cfg


def cases():
    """name -> workload configuration (steps included)."""
    return {
        "rake_4x4x2_30": parity.small_case(),
        "drill_8x8x6_20": drill_small(),
        "translating_8x8x4_45": translating_slab(0.2, 45),
    }


def run_oracle(cfg):
    oracle.set_threads(1)
    o, _ = parity.build_oracle(cfg)
    for _ in range(cfg["steps"]):
        o.step(cfg["dt"])
    p = o.get_particles()
    g = o.grid()
    snap = o.grid_snapshot()
    m = o.metrics()
    out = {f"p_{k}": np.asarray(p[k]) for k in PARTICLE_FIELDS}
    out["g_dims"] = np.array([g.nx, g.ny, g.nz], dtype=np.int64)
    out["g_mass"] = snap["mass"]
    out["g_mom"] = snap["mom"]
    out["metrics"] = np.array([m.max_stress, m.max_temperature, m.kinetic_energy, m.contact_force])
    return out


def path(name):
    return os.path.join(HERE, f"{name}.npz")


def main():
    for name, cfg in cases().items():
        out = run_oracle(cfg)
        np.savez_compressed(path(name), **out)
        print(f"{name}: {len(out['p_id'])} particles, grid {tuple(out['g_dims'])} -> {path(name)}")


if __name__ == "__main__":
    main()
