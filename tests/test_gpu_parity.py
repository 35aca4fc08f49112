"""GPU parity: the HIP solver (through the C ABI) against the CPU FP64 oracle.

Tolerances (normwise relative error, see parity.normwise):
  * FP64 solver vs FP64 oracle: 1e-9 on every compared field (differences
    come only from float-atomic summation order, FMA contraction and libm
    last-bit differences);
  * FP32 solver vs FP64 oracle: 1e-5 on grid mass, grid momentum, particle
    positions, plastic strain and temperature (the gated fields); the other
    fields are reported with looser stated bounds.
"""
import math

import numpy as np
import pytest

import parity
import workloads

pytestmark = pytest.mark.gpu

GATED = ("pos", "temperature", "plastic_strain", "grid_mass", "grid_momentum")


def _report(tag, res):
    e = res["errors"]
    print(f"\n[{tag}] n={res['n']} status_mismatch={res['status_mismatch']} "
          + " ".join(f"{k}={v:.2e}" for k, v in sorted(e.items())), flush=True)


def test_small_fp64_matches_oracle():
    res = parity.run_case(parity.small_case(), "fp64", nsteps=30)
    _report("small fp64", res)
    assert res["status_mismatch"] == 0 and res["grid_same"]
    for k, v in res["errors"].items():
        assert v <= 1e-9, (k, v)


def test_small_fp32_matches_oracle():
    res = parity.run_case(parity.small_case(), "fp32", nsteps=30)
    _report("small fp32", res)
    assert res["status_mismatch"] == 0 and res["grid_same"]
    for k in GATED:
        assert res["errors"][k] <= 1e-5, (k, res["errors"][k])
    assert res["errors"]["stress"] <= 1e-3


@pytest.fixture(scope="module")
def config0_trajectory():
    cfg = workloads.slab_config0()
    states, _ = parity.oracle_trajectory(cfg, {0, 1, 24, 60, 120, 199})
    return cfg, states


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_config0_step_parity_along_trajectory(config0_trajectory, prec, tol):
    """Two steps (standalone P2G, then the fused G2P+P2G path) from oracle
    states taken along the 200-step config-0 trajectory: the GPU step must
    reproduce the reference step on every gated field."""
    cfg, states = config0_trajectory
    for k, st in sorted(states.items()):
        res = parity.run_from_state(cfg, prec, st, 2)
        _report(f"config0 {prec} from step {k}", res)
        assert res["status_mismatch"] == 0, k
        for f in GATED:
            assert res["errors"][f] <= tol, (k, f, res["errors"][f])


@pytest.fixture(scope="module")
def config0_ensemble():
    """The oracle's own sensitivity on the full 200 steps: the same input in
    reversed particle order (another summation order, FP64-rounding-sized
    perturbations), and with initial positions perturbed at FP32 rounding
    size.  The free-surface particles next to near-empty nodes make this run
    chaotic (MPMWorkpieceSolver.cu:160-162 divides node force by node mass)."""
    cfg = workloads.slab_config0()
    n = 32 * 32 * 16 * 8
    last = cfg["steps"]
    ref, _ = parity.oracle_trajectory(cfg, {last})
    rev, _ = parity.oracle_trajectory(cfg, {last}, order=np.arange(n)[::-1])
    pert, _ = parity.oracle_trajectory(cfg, {last}, perturb=6e-8, seed=3)
    return cfg, ref[last], rev[last], pert[last]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_config0_200_steps_within_reference_spread(config0_ensemble, prec):
    cfg, ref, rev, pert = config0_ensemble
    s = parity.build_gpu(cfg, prec)
    for _ in range(cfg["steps"]):
        s.step(cfg["dt"])
    gpu = s.get_particles()
    e_gpu, m_gpu = parity.compare_particles(gpu, ref)
    spread = parity.compare_particles(rev if prec == "fp64" else pert, ref)[0]
    floor = 1e-9 if prec == "fp64" else 1e-5
    m_spread = parity.compare_particles(rev if prec == "fp64" else pert, ref)[1]
    print(f"\n[config0 200 {prec}] status_mismatch={m_gpu} (oracle spread {m_spread}) " + " ".join(
        f"{k}={e_gpu[k]:.2e}/spread {spread[k]:.2e}" for k in sorted(e_gpu)), flush=True)
    for f in ("pos", "temperature", "plastic_strain", "vel"):
        assert e_gpu[f] <= max(10.0 * spread[f], floor), (f, e_gpu[f], spread[f])
    # particles whose status differs at the end: no more than the oracle's
    # own perturbed run shows (statuses change at thresholds of eps_p, D, T)
    assert m_gpu <= max(2 * m_spread, 4), (m_gpu, m_spread)
    m = s.metrics()
    assert m.step == 200


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_config0_shadow_oracle(prec, tol):
    """Step parity along the GPU's own 200-step config-0 trajectory: before
    every step the oracle takes the GPU's particle state (its grid, step
    count and rebuild phase are the GPU's), both take the step, and every
    gated particle field and every status must agree.  The end-to-end test
    below compares the two trajectories, which the free surface makes
    chaotic; this one holds every one of the 200 steps to the tolerance."""
    cfg = workloads.slab_config0()
    per_step, worst = parity.shadow_run(cfg, prec)
    mism = {k: m for k, (e, m) in per_step.items() if m}
    print(f"\n[config0 shadow oracle {prec}] steps with a status mismatch: {mism} worst "
          + " ".join(f"{k}={v:.2e}" for k, v in sorted(worst.items())), flush=True)
    assert not mism
    for k, (e, m) in per_step.items():
        for f in worst:
            assert e[f] <= tol, (k, f, e[f])


def _mixed_status(d):
    rng = np.random.default_rng(11)
    n = len(d["mass"])
    d["status"][rng.random(n) < 0.05] = parity.oracle.CHIP
    d["status"][rng.random(n) < 0.03] = parity.oracle.BOUNDARY
    d["status"][rng.random(n) < 0.03] = parity.oracle.INACTIVE
    chips = d["status"] == parity.oracle.CHIP
    d["damage"][chips] = rng.uniform(1.0, 2.0, chips.sum())     # some fracture kicks
    d["stress"][:] = rng.normal(size=(n, 6)) * 2e8
    d["plastic_strain"][:] = rng.uniform(0, 0.1, n)
    d["damage"][~chips] = rng.uniform(0, 0.95, (~chips).sum())   # some in the CDM window
    d["temperature"][:] = rng.uniform(293, 460, n)
    d["fext"][:] = rng.normal(size=(n, 3)) * 1e-4
    d["residual_stress"][:] = rng.uniform(0, 1e8, n)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mixed_status_chip_fracture_external_force(prec):
    cfg = workloads.scaled(workloads.slab_config0(), (6, 6, 4), steps=26, name="mixed")
    cfg["tip"] = (0.0, 0.0, 1.2e-3)      # chip separation below z = 1.2 mm - 1.5h
    res = parity.run_case(cfg, prec, mutate=_mixed_status)
    _report(f"mixed {prec}", res)
    # step parity along the GPU's own trajectory: every status agrees per step
    d = parity.build_oracle(dict(cfg, mutate=_mixed_status))[1]
    per_step, worst = parity.shadow_run(cfg, prec, particles=d)
    mism = {k: m for k, (e, m) in per_step.items() if m}
    print(f"  shadow oracle: steps with a status mismatch {mism} worst "
          + " ".join(f"{k}={v:.2e}" for k, v in sorted(worst.items())), flush=True)
    tol = 1e-9 if prec == "fp64" else 1e-5
    # every step of the GPU's own trajectory takes the reference step's
    # statuses and values; over the 26 steps the FP32 atomics' order may
    # still move a particle across a status threshold (up to 2 allowed)
    assert not mism
    for k, v in worst.items():
        assert v <= tol, ("shadow", k, v)
    assert res["status_mismatch"] <= (0 if prec == "fp64" else 2)
    for k in GATED:
        assert res["errors"][k] <= tol, (k, res["errors"][k])
    if prec == "fp64":
        for k, v in res["errors"].items():
            assert v <= 1e-8, (k, v)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_kinematic_hardening_backstress(prec):
    cfg = workloads.scaled(workloads.slab_config0(), (6, 6, 4), steps=25, name="kinhard")
    cfg["params"] = dict(cfg["params"], kinematic_hardening=5e9, jc_A=2e8, jc_B=1e8)
    cfg["vel"] = (-3.0, 0.0, 0.0)
    res = parity.run_case(cfg, prec)
    _report(f"kinhard {prec}", res)
    tol = 1e-9 if prec == "fp64" else 1e-5
    assert res["oracle"].get_particles()["plastic_strain"].max() > 0   # yielding happened
    for k in GATED:
        # 25 steps of continuous plastic flow: each FP32 increment
        # (q - sigma_f)/3G is a difference of nearly equal stresses, so the
        # FP32 trajectory accumulates ~1e-5 in eps_p; the step itself is
        # gated at 1e-5 below.
        bound = 1e-4 if (prec == "fp32" and k == "plastic_strain") else tol
        assert res["errors"][k] <= bound, (k, res["errors"][k])
    states, _ = parity.oracle_trajectory(cfg, {0, 12, 24})
    for k, st in sorted(states.items()):
        r = parity.run_from_state(cfg, prec, st, 2)
        _report(f"kinhard {prec} from step {k}", r)
        for f in GATED:
            assert r["errors"][f] <= tol, (k, f, r["errors"][f])


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_drill_tool_small(prec):
    cfg = workloads.scaled(workloads.drill_config1(), (48, 48, 16), steps=20, name="drill_small")
    res = parity.run_case(cfg, prec)
    _report(f"drill {prec}", res)
    tol = 1e-9 if prec == "fp64" else 1e-5
    gm, om = res["metrics"]
    assert res["status_mismatch"] == 0
    for k in GATED:
        assert res["errors"][k] <= tol, (k, res["errors"][k])


def test_empty_and_all_inactive():
    import mpmw
    s = mpmw.Solver(**workloads.params_dict())
    s.step(1e-8)                                  # no particles: silent return
    assert s.metrics().step == 0
    d = mpmw.make_arrays(16)
    d["pos"][:] = np.random.default_rng(0).uniform(0, 1e-3, (16, 3))
    d["mass"][:] = 1e-9
    d["volume"][:] = 2e-12
    d["temperature"][:] = 293.0
    d["status"][:] = mpmw.INACTIVE
    d["id"][:] = np.arange(16)
    s.set_particles(d)
    for _ in range(3):
        s.step(1e-8)
    o = s.get_particles()
    assert np.array_equal(o["pos"], d["pos"])      # positions are FP64 in every mode
    s.compact()
    assert s.count() == 16                         # nothing survives: left as is (:502-506)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_thermal_coupling_api(prec):
    cfg = workloads.scaled(workloads.slab_config0(), (4, 4, 2), steps=5, name="thermal_api")
    o, d = parity.build_oracle(cfg)
    s = parity.build_gpu(cfg, prec)
    q = d["pos"][17] + 1e-6
    assert s.temperature_at(*q) == pytest.approx(o.temperature_at(*q), rel=1e-6)
    s.apply_heat_flux(*q, 1e-3)
    o.apply_heat_flux(*q, 1e-3)
    T = np.linspace(300, 400, 20)
    gd = s.get_particles()
    # device order may differ from the oracle's: set through the device order
    # and give the oracle the same temperatures by id
    s.set_particle_temperatures(T)
    ids = gd["id"][:20]
    od = o.get_particles()
    od["temperature"][np.searchsorted(od["id"], ids)] = T
    o.set_particles(od)
    for _ in range(5):
        s.step(1e-8)
        o.step(1e-8)
    errs, mism = parity.compare_particles(s.get_particles(), o.get_particles())
    tol = 1e-9 if prec == "fp64" else 1e-5
    assert mism == 0 and errs["temperature"] <= tol
    s.sync_metrics()
    o.sync_metrics()
    assert s.metrics().max_temperature == pytest.approx(o.metrics().max_temperature, rel=tol * 10)
    b, ob = s.bounds(), o.bounds()
    assert np.allclose(b, ob, rtol=1e-6, atol=1e-9)


def test_config1_full_size_invariants():
    """16.8M particles, drill tool: mass conservation, finiteness, tool motion."""
    import mpmw
    cfg = workloads.drill_config1()
    P = dict(cfg["params"], precision=mpmw.FP32)
    s = mpmw.Solver(**P)
    workloads.apply_to_solver(mpmw, s, cfg)
    s.set_grid_snapshot(True)
    assert s.count() == 256 * 128 * 64 * 8
    for _ in range(30):
        s.step(cfg["dt"])
    g = s.grid_snapshot()
    d = s.get_particles()
    live = d["status"] != mpmw.INACTIVE
    assert g["mass"].sum() == pytest.approx(d["mass"][live].sum(), rel=1e-4)
    for k in ("pos", "vel", "stress", "temperature"):
        assert np.isfinite(d[k]).all(), k
    m = s.metrics()
    assert m.step == 30 and m.contact_force > 0.0


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_device_slab_matches_host_generator(prec):
    """init_jittered_slab generates the slab on the device; positions, ids,
    statuses, velocities and mass must equal the host generator (itself
    pinned bit for bit against the oracle's)."""
    import mpmw
    cfg = workloads.scaled(workloads.drill_config1(), (13, 7, 9), steps=1)
    P = dict(cfg["params"], precision=mpmw.FP64 if prec == "fp64" else mpmw.FP32)
    s = mpmw.Solver(**P)
    vel = (0.5, -0.25, 0.125)
    s.init_jittered_slab(cfg["seed"], 13, 7, 9, (1e-3, -2e-3, 5e-4), vel, 301.0)
    g = s.get_particles()
    h = mpmw.gen_jittered_slab(cfg["seed"], 13, 7, 9, workloads.H, (1e-3, -2e-3, 5e-4), P["density"],
                               P["fixture_thickness"])
    assert np.array_equal(g["pos"], h["pos"])
    assert np.array_equal(g["id"], h["id"]) and np.array_equal(g["status"], h["status"])
    fixed = h["status"] == mpmw.FIXED_BOUNDARY
    assert fixed.any() and (~fixed).any()
    assert np.all(g["vel"][fixed] == 0.0)
    want = np.array(vel, dtype=np.float32 if prec == "fp32" else np.float64).astype(np.float64)
    assert np.all(g["vel"][~fixed] == want)
    assert np.allclose(g["mass"], h["mass"], rtol=1e-7) and np.all(g["temperature"] == 301.0)
    s.close()
