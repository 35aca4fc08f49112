"""GPU-vs-oracle parity harness shared by the GPU tests and smoke().

Runs one workload on the HIP solver (through the C ABI) and on the CPU FP64
oracle from identical seeded inputs, then compares particle fields (matched
by id) and the post-P2G grid of the last step.

Error measure per field: normwise relative error
    max_i |gpu_i - oracle_i| / max_i |oracle_i|
(zero when both fields are identically zero), the tolerance the tests state
is on this number.
"""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "edegpredict-engine-rotary_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import mpmw  # noqa: E402
import oracle  # noqa: E402
import workloads  # noqa: E402

FIELDS = ("pos", "vel", "temperature", "plastic_strain", "stress", "damage", "strain_rate",
          "residual_stress")
ORACLE_PARAM_FIELDS = {f for f, _ in oracle.Params._fields_}


def small_case(**over):
    """4x4x2-cell slab (256 particles) against the rake-face point tool."""
    cfg = workloads.scaled(workloads.slab_config0(), (4, 4, 2), steps=30, name="small_rake_4x4x2")
    cfg = dict(cfg)
    cfg.update(over)
    return cfg


def build_oracle(cfg):
    P = cfg["params"]
    o = oracle.Oracle(**{k: v for k, v in P.items() if k in ORACLE_PARAM_FIELDS})
    cx, cy, cz = cfg["cells"]
    fixture = P["fixture_thickness"] if P.get("has_config", 1) else 0.0
    d = oracle.gen_jittered_slab(cfg["seed"], cx, cy, cz, P["cell_size"], cfg["origin"], P["density"], fixture)
    live = d["status"] != oracle.FIXED_BOUNDARY
    d["vel"][live] = cfg["vel"]
    d["temperature"][:] = cfg["temperature"]
    if cfg.get("mutate"):
        cfg["mutate"](d)
    o.set_particles(d)
    if cfg.get("transfer") == "apic":
        o.set_transfer(oracle.TRANSFER_APIC)
    elif cfg.get("transfer") == "apic_textbook":
        o.set_transfer(oracle.TRANSFER_APIC_TEXTBOOK)
    kind = cfg["tool"][0]
    if kind == "points":
        o.set_tool_samples(cfg["tool"][1], cfg["tool"][2])
    elif kind in ("drill", "wedge", "voxel"):
        desc = oracle.ToolDesc()
        for k, v in cfg["tool"][1].items():
            setattr(desc, k, v)
        if kind == "voxel":
            verts, tris, normals = cfg["tool"][3]
            o.set_tool_sdf(oracle.SdfGrid(verts, tris, normals, cfg["tool"][4], cfg["tool"][5]), desc)
        else:
            o.set_tool_analytic(desc)
        pose = oracle.ToolPose()
        pd = cfg["tool"][2]
        pose.center[:] = pd["center"]
        pose.angle = pd["angle"]
        pose.vel[:] = pd["vel"]
        pose.omega = pd["omega"]
        o.set_tool_pose(pose)
    o.set_tool_tip(*cfg["tip"])
    return o, d


def build_gpu(cfg, precision, deterministic=False, particles=None):
    P = dict(cfg["params"])
    P["precision"] = mpmw.FP64 if precision == "fp64" else mpmw.FP32
    P["deterministic"] = 1 if deterministic else 0
    s = mpmw.Solver(**P)
    workloads.apply_to_solver(mpmw, s, cfg)
    if particles is not None:
        s.set_particles(particles)
    return s


def normwise(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max() if b.size else 0.0
    num = np.abs(a - b).max() if b.size else 0.0
    if den == 0.0:
        return 0.0 if num == 0.0 else math.inf
    return float(num / den)


def compare_particles(g, o):
    gi = np.argsort(g["id"], kind="stable")
    oi = np.argsort(o["id"], kind="stable")
    assert np.array_equal(g["id"][gi], o["id"][oi]), "particle id sets differ"
    same = g["status"][gi] == o["status"][oi]
    errs = {}
    for f in FIELDS:
        errs[f] = normwise(g[f][gi][same], o[f][oi][same])
    return errs, int((~same).sum())


def oracle_trajectory(cfg, checkpoints, order=None, perturb=0.0, seed=0, perturb_vel=0.0, fp32_emulation=False):
    """Run the oracle over cfg['steps'] steps; return {step: particle dict}
    for every step index in `checkpoints` (state before that step).  `order`
    permutes the particle input order (changes the P2G summation order);
    `perturb` adds relative noise of that size to the initial positions,
    `perturb_vel` to the initial velocities; `fp32_emulation` rounds the
    P2G node sums after every add and the stored particle state (all but
    the positions) after every step to FP32, as an FP32 solver keeps them."""
    o, d = build_oracle(cfg)
    if fp32_emulation:
        o.set_fp32_emulation(True)
    if order is not None or perturb or perturb_vel:
        d = {k: (None if v is None else v.copy()) for k, v in d.items()}
        if perturb or perturb_vel:
            rng = np.random.default_rng(seed)
            d["pos"] *= 1.0 + perturb * rng.uniform(-1, 1, d["pos"].shape)
            d["vel"] *= 1.0 + perturb_vel * rng.uniform(-1, 1, d["vel"].shape)
        if order is not None:
            for k, v in d.items():
                if v is not None:
                    d[k] = np.ascontiguousarray(v[order])
        o.set_particles(d)
    out = {}
    for i in range(cfg["steps"] + 1):
        if i in checkpoints:
            out[i] = o.get_particles()
        if i < cfg["steps"]:
            o.step(cfg["dt"])
    return out, o


def run_from_state(cfg, precision, state, nsteps, deterministic=False):
    """Load the same particle state into a fresh oracle and a fresh GPU
    solver and step both `nsteps` times (step-function parity)."""
    o, _ = build_oracle(cfg)
    o.set_particles(state)
    s = build_gpu(cfg, precision, deterministic, particles=state)
    s.set_grid_snapshot(True)
    for _ in range(nsteps):
        o.step(cfg["dt"])
        s.step(cfg["dt"])
    errs, mism = compare_particles(s.get_particles(), o.get_particles())
    gg, og = s.grid_info(), o.grid()
    if (gg.nx, gg.ny, gg.nz) == (og.nx, og.ny, og.nz):
        gs, osn = s.grid_snapshot(), o.grid_snapshot()
        errs["grid_mass"] = normwise(gs["mass"], osn["mass"])
        errs["grid_momentum"] = normwise(gs["mom"], osn["mom"])
        errs["grid_force"] = normwise(gs["force"], osn["force"])
        errs["grid_tm"] = normwise(gs["tm"], osn["tm"])
    else:
        errs["grid_mass"] = math.inf
    return {"errors": errs, "status_mismatch": mism, "n": len(state["mass"])}


def run_case(cfg, precision, nsteps=None, deterministic=False, snapshot=True, mutate=None):
    """Step the GPU solver and the oracle side by side; return errors."""
    if mutate is not None:
        cfg = dict(cfg)
        cfg["mutate"] = mutate
    nsteps = cfg["steps"] if nsteps is None else nsteps
    o, d = build_oracle(cfg)
    s = build_gpu(cfg, precision, deterministic, particles=d if cfg.get("mutate") else None)
    if snapshot:
        s.set_grid_snapshot(True)
    dt = cfg["dt"]
    for _ in range(nsteps):
        o.step(dt)
        s.step(dt)
    gp = s.get_particles()
    op = o.get_particles()
    errs, status_mismatch = compare_particles(gp, op)
    out = {"errors": errs, "status_mismatch": status_mismatch, "n": len(op["mass"])}
    if snapshot and nsteps > 0:
        gg, og = s.grid_info(), o.grid()
        same_grid = (gg.nx, gg.ny, gg.nz) == (og.nx, og.ny, og.nz)
        out["grid_same"] = same_grid
        if same_grid:
            gs, osn = s.grid_snapshot(), o.grid_snapshot()
            errs["grid_mass"] = normwise(gs["mass"], osn["mass"])
            errs["grid_momentum"] = normwise(gs["mom"], osn["mom"])
            errs["grid_force"] = normwise(gs["force"], osn["force"])
            errs["grid_tm"] = normwise(gs["tm"], osn["tm"])
        else:
            errs["grid_mass"] = math.inf
    gm, om = s.metrics(), o.metrics()
    out["metrics"] = (gm, om)
    out["solver"] = s
    out["oracle"] = o
    return out


def shadow_run(cfg, precision, nsteps=None, deterministic=False, fields=("pos", "temperature", "plastic_strain",
                                                                        "vel"), particles=None):
    """Step parity along the GPU's own trajectory: before every step the
    oracle takes the GPU's particle state (orc_replace_particles keeps the
    oracle's grid, step count and rebuild phase, which follow the same rule
    as the GPU's), then both take the step.  Returns the per-step errors
    {step: (errors, status_mismatch)} and the worst error per field."""
    nsteps = cfg["steps"] if nsteps is None else nsteps
    o, d = build_oracle(cfg)
    s = build_gpu(cfg, precision, deterministic, particles=particles)
    if particles is not None:
        o.set_particles(particles)
    per_step, worst = {}, {f: 0.0 for f in fields}
    for k in range(nsteps):
        o.replace_particles(s.get_particles())
        o.step(cfg["dt"])
        s.step(cfg["dt"])
        g, og = s.grid_info(), o.grid()
        assert (g.nx, g.ny, g.nz) == (og.nx, og.ny, og.nz), ("grid differs", k)
        e, m = compare_particles(s.get_particles(), o.get_particles())
        per_step[k] = (e, m)
        for f in fields:
            worst[f] = max(worst[f], e[f])
    s.close()
    return per_step, worst
