"""APIC transfer variant on the device against the oracle and a known answer.

* Oracle parity (FP64 1e-9, FP32 1e-5 normwise on the gated fields, the
  post-P2G grid and the affine matrices): the rake-face point tool on a small
  slab over 60 steps (two grid rebuilds, so the G2P-only pass, re-binning
  and the P2G after a rebuild run), the analytic wedge of the batched
  distribution, and the same wedge workpiece inside a batch with trilinear
  members, and against the insert mesh's voxel SDF.
* Known answer without the oracle: an affine velocity field v0 + A x with
  C = A gives C = 0.7425 A and v = v0 + A x after one step; with the
  textbook transfer (4/h^2, no decay) C = A.
* The textbook transfer against the oracle on the rake case.
* The deterministic mode is refused with APIC.
"""
import numpy as np
import pytest

import mpmw
import oracle
import parity
import workloads

pytestmark = pytest.mark.gpu

GATED = ("pos", "temperature", "plastic_strain")
H = workloads.H


def affine_errors(res):
    s, o = res["solver"], res["oracle"]
    g, op = s.get_particles(), o.get_particles()
    gi, oi = np.argsort(g["id"], kind="stable"), np.argsort(op["id"], kind="stable")
    return parity.normwise(s.get_affine()[gi], o.get_affine()[oi])


def check(res, tol, tag):
    e = res["errors"]
    e["affine"] = affine_errors(res)
    gm, om = res["metrics"]
    print(f"\n[apic {tag}] cf gpu/oracle {gm.contact_force:.6g}/{om.contact_force:.6g} "
          + " ".join(f"{k}={v:.2e}" for k, v in sorted(e.items())), flush=True)
    assert res["status_mismatch"] == 0 and res["grid_same"]
    for k in GATED + ("grid_mass", "grid_momentum", "grid_force", "affine"):
        assert e[k] <= tol, (k, e[k])


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_apic_rake_matches_oracle(prec, tol):
    cfg = parity.small_case(steps=60, transfer="apic", name="apic_rake_4x4x2")
    res = parity.run_case(cfg, prec)
    check(res, tol, f"rake {prec}")
    assert np.abs(res["solver"].get_affine()).max() > 0.0


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_apic_textbook_rake_matches_oracle(prec, tol):
    cfg = parity.small_case(steps=60, transfer="apic_textbook", name="apic_textbook_rake_4x4x2")
    res = parity.run_case(cfg, prec)
    check(res, tol, f"textbook rake {prec}")
    assert np.abs(res["solver"].get_affine()).max() > 0.0


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_apic_wedge_matches_oracle(prec, tol):
    cfg = sorted(workloads.batch_config2(), key=lambda c: np.prod(c["cells"]))[0]
    cfg = dict(cfg, steps=30, transfer="apic")
    res = parity.run_case(cfg, prec)
    check(res, tol, f"wedge {cfg['name']} {prec}")


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_apic_with_voxel_tool_matches_oracle(prec, tol):
    """The APIC transfer against the insert mesh's voxel SDF."""
    cfg = sorted(workloads.batch_config2(), key=lambda c: np.prod(c["cells"]))[0]
    cfg = dict(workloads.voxel_tool_case(dict(cfg, steps=30)), transfer="apic")
    res = parity.run_case(cfg, prec)
    assert res["metrics"][1].contact_force > 0.0
    check(res, tol, f"voxel {cfg['name']} {prec}")


def test_apic_member_of_mixed_batch():
    cfgs = sorted(workloads.batch_config2(), key=lambda c: np.prod(c["cells"]))[:3]
    cfgs = [dict(c, steps=30) for c in cfgs]
    cfgs[1] = dict(cfgs[1], transfer="apic")
    solvers = [parity.build_gpu(c, "fp64") for c in cfgs]
    b = mpmw.Batch()
    for s in solvers:
        b.add(s)
    for _ in range(30):
        b.step([c["dt"] for c in cfgs])
    b.synchronize()
    o, _ = parity.build_oracle(cfgs[1])
    for _ in range(30):
        o.step(cfgs[1]["dt"])
    errs, mism = parity.compare_particles(solvers[1].get_particles(), o.get_particles())
    print("\n[apic in batch] " + " ".join(f"{k}={v:.2e}" for k, v in sorted(errs.items())), flush=True)
    assert mism == 0
    for k in GATED:
        assert errs[k] <= 1e-9, (k, errs[k])
    b.close()
    for s in solvers:
        s.close()


@pytest.mark.parametrize("mode,factor", [(mpmw.TRANSFER_APIC, 0.7425), (mpmw.TRANSFER_APIC_TEXTBOOK, 1.0)])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 5e-5)])
def test_apic_affine_field_known_answer(prec, tol, mode, factor):
    """FP32 bound: B = sum w v_i (x_i - x_p)^T cancels the offset v0 out of
    node velocities of ~0.4 m/s whose FP32 rounding (~3e-8 m/s) stands against
    |A| h / 4 ~ 2.5e-3 m/s per node, ~1e-5 of C per node over 27 nodes."""
    cfg = workloads.scaled(workloads.slab_config0(), (6, 6, 4), steps=1)
    P = dict(cfg["params"], fixture_thickness=0.0, precision=mpmw.FP64 if prec == "fp64" else mpmw.FP32)
    s = mpmw.Solver(**P)
    d = oracle.gen_jittered_slab(cfg["seed"], 6, 6, 4, H, (0, 0, 0), P["density"], 0.0)
    A = np.array([[40.0, -25.0, 10.0], [15.0, -30.0, 5.0], [-20.0, 35.0, -10.0]])
    v0 = np.array([0.3, -0.2, 0.1])
    x0 = d["pos"].copy()
    d["vel"][:] = v0 + x0 @ A.T
    d["temperature"][:] = 293.0
    s.set_particles(d)
    s.set_transfer(mode)
    s.set_affine(np.tile(A.reshape(9), (len(x0), 1)))
    s.step(1e-9)
    out = s.get_particles()
    i = np.argsort(out["id"])
    C9 = s.get_affine()[i].reshape(-1, 3, 3)
    v = out["vel"][i]
    ref_v = v0 + x0 @ A.T
    assert np.abs(C9 - factor * A).max() <= tol * np.abs(A).max()
    assert np.abs(v - ref_v).max() <= tol * np.abs(ref_v).max()
    s.close()


def test_apic_refuses_deterministic():
    cfg = parity.small_case(steps=2)
    s = parity.build_gpu(cfg, "fp32", deterministic=True)
    with pytest.raises(mpmw.MPMError):
        s.set_transfer(mpmw.TRANSFER_APIC)
    s.set_deterministic(False)
    s.set_transfer(mpmw.TRANSFER_APIC)
    with pytest.raises(mpmw.MPMError):
        s.set_deterministic(True)
    s.step(cfg["dt"])
    s.set_transfer(mpmw.TRANSFER_PIC)
    with pytest.raises(mpmw.MPMError):
        s.get_affine()
    s.step(cfg["dt"])
    s.close()
