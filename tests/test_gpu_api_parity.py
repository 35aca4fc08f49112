"""GPU-vs-oracle parity of the solver API around the step.

Each case drives the HIP solver through the C ABI and the CPU FP64 oracle
with the same calls and compares them (particles matched by id):

  * metrics (updateMetrics, MPMWorkpieceSolver.cu:855-868): max von Mises
    stress, max temperature and kinetic energy every 25 steps, contact force;
  * compaction (compactParticles, :464-527) of a mixed ACTIVE / CHIP /
    BOUNDARY / INACTIVE set carrying backstress, external forces and APIC
    affine matrices, by call and at the step-1000 trigger (:787-789);
  * initializeParticleBox (:588-614) and initializeCylindricalWorkpiece
    (:616-648) lattices, then stepped;
  * setToolMesh's stride rule (:793-810) with more than 8000 mesh nodes;
  * applyExternalForce (:812-818), consumed by the next P2G.

Tolerances: FP64 1e-9 normwise; FP32 1e-5 on the gated fields (grid mass
and momentum, positions, plastic strain, temperature) and on kinetic energy;
FP32 max stress 1e-4 (the stress of a yielding particle is the flow stress
times q_trial / q, a ratio of nearly equal FP32 numbers).
"""
import numpy as np
import pytest

import mpmw
import parity
import workloads

pytestmark = pytest.mark.gpu

GATED = ("pos", "temperature", "plastic_strain", "grid_mass", "grid_momentum")
TOL = {"fp64": 1e-9, "fp32": 1e-5}


def _report(tag, e, extra=""):
    print(f"\n[{tag}] {extra} " + " ".join(f"{k}={v:.2e}" for k, v in sorted(e.items())), flush=True)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def _by_id(d):
    return np.argsort(d["id"], kind="stable")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_metrics_match_oracle(prec):
    """Max von Mises stress, max temperature and kinetic energy of the
    non-INACTIVE particles after 25 and 50 steps (the refresh steps), and the
    last step's contact force, against the oracle's."""
    cfg = workloads.scaled(workloads.slab_config0(), (6, 6, 4), steps=50, name="metrics")
    o, _ = parity.build_oracle(cfg)
    s = parity.build_gpu(cfg, prec)
    for step in range(1, 51):
        o.step(cfg["dt"])
        s.step(cfg["dt"])
        if step % 25:
            continue
        gm, om = s.metrics(), o.metrics()
        assert gm.step == om.step == step
        e = {f: _rel(getattr(gm, f), getattr(om, f))
             for f in ("max_stress", "max_temperature", "kinetic_energy", "contact_force")}
        _report(f"metrics {prec} step {step}", e)
        assert om.max_stress > 1e6 and om.kinetic_energy > 0 and om.contact_force > 0
        tol = TOL[prec]
        assert e["max_temperature"] <= tol and e["kinetic_energy"] <= tol, e
        assert e["max_stress"] <= (tol if prec == "fp64" else 1e-4), e
        assert e["contact_force"] <= (tol if prec == "fp64" else 1e-4), e


def _mixed(d, seed=21):
    rng = np.random.default_rng(seed)
    n = len(d["mass"])
    r = rng.random(n)
    d["status"][r < 0.08] = parity.oracle.CHIP
    d["status"][(r >= 0.08) & (r < 0.12)] = parity.oracle.BOUNDARY
    d["status"][(r >= 0.12) & (r < 0.22)] = parity.oracle.INACTIVE
    d["stress"][:] = rng.normal(size=(n, 6)) * 1e8
    d["backstress"][:] = rng.normal(size=(n, 6)) * 1e7
    d["plastic_strain"][:] = rng.uniform(0, 0.05, n)
    d["temperature"][:] = rng.uniform(293, 400, n)
    d["fext"][:] = rng.normal(size=(n, 3)) * 1e-4
    d["residual_stress"][:] = rng.uniform(0, 1e8, n)


def _compare_all(g, o, fields=parity.FIELDS + ("mass", "volume", "fext", "backstress", "status")):
    gi, oi = _by_id(g), _by_id(o)
    assert np.array_equal(g["id"][gi], o["id"][oi]), "particle id sets differ"
    return {f: parity.normwise(g[f][gi], o[f][oi]) for f in fields if g.get(f) is not None}


@pytest.mark.parametrize("transfer", ["pic", "apic"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_compact_mixed_set_matches_oracle(prec, transfer):
    """compact() of a mixed set with kinematic hardening (backstress
    gathered), pending external forces and, with APIC, the affine matrices:
    the survivors are exactly the non-INACTIVE particles, every field of
    every survivor equals the oracle's, and stepping on after the compaction
    stays on the oracle."""
    cfg = workloads.scaled(workloads.slab_config0(), (6, 6, 4), steps=10, name="compact_mixed")
    cfg["params"] = dict(cfg["params"], kinematic_hardening=3e9)
    cfg["tip"] = (0.0, 0.0, -1.0)
    if transfer == "apic":
        cfg["transfer"] = "apic"
    cfg["mutate"] = _mixed
    o, d = parity.build_oracle(cfg)
    s = parity.build_gpu(cfg, prec, particles=d)
    for _ in range(4):
        o.step(cfg["dt"])
        s.step(cfg["dt"])
    n0 = s.count()
    live = int((o.get_particles()["status"] != parity.oracle.INACTIVE).sum())
    assert 0 < live < n0
    # pending external forces travel with their particles through the compaction
    for q in (o, s):
        q.apply_external_force((0.75e-3, 0.75e-3, 0.5e-3), 0.6e-3, (1e-3, -2e-3, 5e-4))
    o.compact()
    s.compact()
    assert s.count() == o.count() == live
    g, op = s.get_particles(), o.get_particles()
    assert not (g["status"] == mpmw.INACTIVE).any()
    e = _compare_all(g, op)
    assert np.abs(op["fext"]).max() > 0 and np.abs(op["backstress"]).max() > 0
    if transfer == "apic":
        e["affine"] = parity.normwise(s.get_affine()[_by_id(g)], o.get_affine()[_by_id(op)])
    _report(f"compact {transfer} {prec}", e, f"survivors={live} of {n0}")
    tol = TOL[prec]
    for k, v in e.items():
        assert v <= (tol if prec == "fp64" or k in ("pos", "status", "mass", "volume") else 1e-4), (k, v)
    assert e["status"] == 0.0
    for _ in range(6):
        o.step(cfg["dt"])
        s.step(cfg["dt"])
    errs, mism = parity.compare_particles(s.get_particles(), o.get_particles())
    _report(f"compact {transfer} {prec} then 6 steps", errs, f"status_mismatch={mism}")
    assert mism == 0
    for k in ("pos", "temperature", "plastic_strain"):
        assert errs[k] <= tol, (k, errs[k])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_compaction_trigger_at_step_1000(prec):
    """The step function compacts after its 1000th step (:787-789): one step
    before, INACTIVE particles are still held; after it, exactly the oracle's
    survivors remain, with the oracle's fields."""
    cfg = workloads.scaled(workloads.slab_config0(), (3, 3, 2), steps=1000, name="compact_1000")
    cfg["tool"] = ("none",)
    cfg["tip"] = (0.0, 0.0, -1.0)
    cfg["vel"] = (0.3, -0.2, 0.1)

    def some_inactive(d):
        d["status"][::7] = parity.oracle.INACTIVE

    cfg["mutate"] = some_inactive
    o, d = parity.build_oracle(cfg)
    s = parity.build_gpu(cfg, prec, particles=d)
    n0 = len(d["mass"])
    inactive = int((d["status"] == parity.oracle.INACTIVE).sum())
    s.run(999, cfg["dt"])
    o.run(999, cfg["dt"])
    assert s.count() == o.count() == n0
    s.step(cfg["dt"])
    o.step(cfg["dt"])
    assert s.metrics().step == 1000
    assert s.count() == o.count() == n0 - inactive
    errs, mism = parity.compare_particles(s.get_particles(), o.get_particles())
    _report(f"compact at 1000 {prec}", errs, f"status_mismatch={mism}")
    assert mism == 0
    for k in ("pos", "temperature", "plastic_strain", "vel"):
        assert errs[k] <= TOL[prec], (k, errs[k])


def _lattice_same(s, o, prec):
    g, op = s.get_particles(), o.get_particles()
    assert len(g["id"]) == len(op["id"]) > 0
    gi, oi = _by_id(g), _by_id(op)
    assert np.array_equal(g["id"][gi], op["id"][oi])
    assert np.array_equal(g["pos"][gi], op["pos"][oi])          # FP64 positions in every mode
    assert np.array_equal(g["status"][gi], op["status"][oi])
    rt = 0.0 if prec == "fp64" else 1e-7
    for f in ("mass", "volume", "temperature"):
        assert np.allclose(g[f][gi], op[f][oi], rtol=rt, atol=0.0), f
    return g


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_init_particle_box_matches_oracle(prec):
    """initializeParticleBox's lattice (z descending from the top, y and x
    ascending, FIXED_BOUNDARY within the fixture layer), then 10 steps with
    an external push on both.  The lattice puts every particle on a node
    plane or a cell's mid-plane, so the push moves particles off node planes
    by fractions of ~1e-7: nodes that only such tiny weights reach take
    accelerations ~ 1/fraction, and the FP32 run's trajectory is gated by
    step parity (two steps from the oracle's states at steps 0 and 8)
    instead of end to end."""
    P = workloads.params_dict(fixture=0.3e-3)
    o = parity.oracle.Oracle(**{k: v for k, v in P.items() if k in parity.ORACLE_PARAM_FIELDS})
    s = mpmw.Solver(**dict(P, precision=mpmw.FP64 if prec == "fp64" else mpmw.FP32))
    mn, mx, sp = (0.1e-3, -0.2e-3, 0.0), (1.6e-3, 1.1e-3, 1.0e-3), 1.25e-4
    o.init_box(mn, mx, sp)
    s.init_box(mn, mx, sp)
    g = _lattice_same(s, o, prec)
    fixed = g["status"] == mpmw.FIXED_BOUNDARY
    assert fixed.any() and (~fixed).any()
    for q in (o, s):
        q.apply_external_force((0.8e-3, 0.4e-3, 0.9e-3), 0.5e-3, (3e-3, -1e-3, -2e-3))
    if prec == "fp64":
        for _ in range(10):
            o.step(1e-8)
            s.step(1e-8)
        errs, mism = parity.compare_particles(s.get_particles(), o.get_particles())
        _report(f"box {prec}", errs, f"n={len(g['id'])} status_mismatch={mism}")
        assert mism == 0
        for k in ("pos", "temperature", "plastic_strain", "vel"):
            assert errs[k] <= TOL[prec], (k, errs[k])
    else:
        states = {0: o.get_particles()}
        for i in range(8):
            o.step(1e-8)
        states[8] = o.get_particles()
        for k, st in states.items():
            a = parity.oracle.Oracle(**{k2: v for k2, v in P.items() if k2 in parity.ORACLE_PARAM_FIELDS})
            b = mpmw.Solver(**dict(P, precision=mpmw.FP32))
            a.set_particles(st)
            b.set_particles(st)
            for _ in range(2):
                a.step(1e-8)
                b.step(1e-8)
            errs, mism = parity.compare_particles(b.get_particles(), a.get_particles())
            _report(f"box {prec} from step {k}", errs, f"n={len(g['id'])} status_mismatch={mism}")
            assert mism == 0
            for f in ("pos", "temperature", "plastic_strain", "vel"):
                assert errs[f] <= TOL[prec], (k, f, errs[f])
            b.close()
    s.close()


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_init_cylinder_matches_oracle(prec, axis):
    """initializeCylindricalWorkpiece's lattice, then 5 steps.  The reference
    ignores the axis argument (MPMWorkpieceSolver.cu:618, `(void)axis`): the
    cylinder always runs along z, whatever axis is passed."""
    P = workloads.params_dict(fixture=0.2e-3)
    o = parity.oracle.Oracle(**{k: v for k, v in P.items() if k in parity.ORACLE_PARAM_FIELDS})
    s = mpmw.Solver(**dict(P, precision=mpmw.FP64 if prec == "fp64" else mpmw.FP32))
    c, r, L, sp = (1.0e-3, 0.8e-3, 0.6e-3), 0.55e-3, 1.1e-3, 1.25e-4
    o.init_cylinder(c, r, L, sp, axis)
    s.init_cylinder(c, r, L, sp, axis)
    g = _lattice_same(s, o, prec)
    rad = (g["pos"] - np.array(c))[:, :2]
    assert (np.hypot(rad[:, 0], rad[:, 1]) <= r + 1e-12).all()
    assert np.abs(g["pos"][:, 2] - c[2]).max() <= 0.5 * L + 1e-12
    for _ in range(5):
        o.step(1e-8)
        s.step(1e-8)
    errs, mism = parity.compare_particles(s.get_particles(), o.get_particles())
    _report(f"cylinder axis {axis} {prec}", errs, f"n={len(g['id'])} status_mismatch={mism}")
    assert mism == 0
    for k in ("pos", "temperature", "vel"):
        assert errs[k] <= TOL[prec], (k, errs[k])
    s.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_tool_samples_stride_rule(prec):
    """A rake-face patch of 9,409 mesh nodes: setToolMesh keeps every second
    one (stride = 9409 / 4000 = 2) and the contact against the kept subset
    equals the oracle's, contact force included."""
    cfg = workloads.scaled(workloads.slab_config0(), (6, 6, 4), steps=20, name="stride")
    top = 4 * workloads.H
    pts = workloads.rake_face_samples(0.0, 6 * workloads.H, top - 0.5e-3, spacing=workloads.H / 16)
    assert len(pts) > 8000
    cfg["tool"] = ("points", pts, 1.5 * workloads.H)
    res = parity.run_case(cfg, prec)
    gm, om = res["metrics"]
    e = dict(res["errors"])
    e["contact_force"] = _rel(gm.contact_force, om.contact_force)
    _report(f"stride {prec}", e, f"samples={len(pts)} status_mismatch={res['status_mismatch']}")
    assert om.contact_force > 0 and res["status_mismatch"] == 0
    for k in GATED:
        assert res["errors"][k] <= TOL[prec], (k, res["errors"][k])
    assert e["contact_force"] <= (TOL[prec] if prec == "fp64" else 1e-4)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_apply_external_force_matches_oracle(prec):
    """applyExternalForce adds F to the ACTIVE particles within the radius
    (CHIP and BOUNDARY particles get none); the next P2G scatters it as
    f_ext w and G2P clears it.  Applied before the first step, again after
    three steps (twice in a row, so the forces add), then stepped on."""
    cfg = workloads.scaled(workloads.slab_config0(), (6, 6, 4), steps=8, name="ext_force")
    cfg["tool"] = ("none",)
    cfg["tip"] = (0.0, 0.0, -1.0)

    def statuses(d):
        d["status"][::5] = parity.oracle.CHIP
        d["status"][1::9] = parity.oracle.BOUNDARY

    cfg["mutate"] = statuses
    o, d = parity.build_oracle(cfg)
    s = parity.build_gpu(cfg, prec, particles=d)
    s.set_grid_snapshot(True)
    c, r = (0.75e-3, 0.75e-3, 0.8e-3), 0.45e-3
    for q in (o, s):
        q.apply_external_force(c, r, (2e-3, 0.0, -1e-3))
    for i in range(8):
        if i == 3:
            for q in (o, s):
                q.apply_external_force(c, r, (-1e-3, 1.5e-3, 0.0))
                q.apply_external_force((0.3e-3, 0.3e-3, 0.5e-3), 0.3e-3, (0.0, 0.0, 4e-3))
        o.step(cfg["dt"])
        s.step(cfg["dt"])
    errs, mism = parity.compare_particles(s.get_particles(), o.get_particles())
    gs, osn = s.grid_snapshot(), o.grid_snapshot()
    errs["grid_force"] = parity.normwise(gs["force"], osn["force"])
    errs["grid_momentum"] = parity.normwise(gs["mom"], osn["mom"])
    _report(f"external force {prec}", errs, f"status_mismatch={mism}")
    assert mism == 0
    moved = np.abs(o.get_particles()["vel"]).max()
    assert moved > 1e-3                               # the forces did act
    for k in ("pos", "vel", "temperature", "grid_force", "grid_momentum"):
        assert errs[k] <= TOL[prec], (k, errs[k])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_reset_then_rerun_equal(prec):
    """reset() clears the particles, tool and step count and puts the tool
    pose back where it was set: the same set-up run again after a reset
    equals the first run bit for bit."""
    cfg = workloads.batch_config2(count=3)[0]
    P = dict(cfg["params"], precision=mpmw.FP64 if prec == "fp64" else mpmw.FP32, deterministic=1)
    s = mpmw.Solver(**P)

    def run():
        workloads.apply_to_solver(mpmw, s, cfg)
        s.run(30, cfg["dt"])
        s.synchronize()
        return s.get_particles()

    a = run()
    s.reset()
    assert s.count() == 0 and s.metrics().step == 0
    b = run()
    for k, v in a.items():
        if v is not None:
            assert np.array_equal(v, b[k]), k
    s.close()
