"""Pins of the oracle's APIC transfer variant (MPMSolver.cu restated).

* Quadratic weights (getWeightsAndGradients, MPMSolver.cu:20-34): a value
  worked by hand; partition of unity, zero gradient sum, linear
  reproduction sum w (x_i - x_p) = 0, sum (x_i - x_p) grad w^T = I and the
  B-spline inertia sum w (x_i - x_p)(x_i - x_p)^T = h^2/4 I at random points;
  the stencil against the textbook B-spline N(u) evaluated here.
* P2G of one particle: every node's momentum m w (v + C (x_i - x_p)) and
  mass against the textbook B-spline computed here; a FIXED_BOUNDARY
  particle adds mass and no momentum (:393-398).
* A whole step on an affine velocity field v = v0 + A x with C = A: node
  velocities are v0 + A x_i exactly, so the G2P returns v0 + A x_p and
  B = A h^2/4, hence C = 0.99 * 3/h^2 * A h^2/4 = 0.7425 A (the reference's
  3/h^2, :578, and 0.99 decay, :591; the textbook 4/h^2 would give 0.99 A).
  The textbook form (4/h^2, no decay) carries the field exactly: C = A.
"""
import numpy as np
import pytest

import oracle
import workloads

H = workloads.H


def bspline(u):
    a = abs(u)
    if a < 0.5:
        return 0.75 - a * a
    if a < 1.5:
        return 0.5 * (1.5 - a) ** 2
    return 0.0


def grid(n=12, origin=(0.0, 0.0, 0.0)):
    g = oracle.Grid()
    g.cell = H
    g.minx, g.miny, g.minz = origin
    g.nx = g.ny = g.nz = n
    return g


def test_quadratic_weights_by_hand_and_identities():
    g = grid()
    # r = (2.3, 5.7, 4.5): base = floor(r - 0.5) = (1, 5, 4); x weights at d0 = 1.3
    idx, w, gr, dp = oracle.weights_quadratic(g, 2.3 * H, 5.7 * H, 4.5 * H)
    wx = [0.5 * 0.2 ** 2, 0.75 - 0.3 ** 2, 0.5 * 0.8 ** 2]
    assert sum(wx) == pytest.approx(1.0, abs=1e-15)
    # node (i, j, k) is entry 9 i + 3 j + k, x outermost
    assert idx[0] == (4 * g.ny + 5) * g.nx + 1
    assert idx[9] == idx[0] + 1 and idx[3] == idx[0] + g.nx and idx[1] == idx[0] + g.nx * g.ny
    rng = np.random.default_rng(3)
    for _ in range(50):
        p = rng.uniform(3.0, 8.0, 3) * H
        idx, w, gr, dp = oracle.weights_quadratic(g, *p)
        assert (idx >= 0).all()
        assert w.sum() == pytest.approx(1.0, abs=1e-14)
        assert np.abs(gr.sum(axis=0)).max() < 1e-9 / H
        assert np.abs((w[:, None] * dp).sum(axis=0)).max() < 1e-14 * H
        assert (dp.T @ gr) == pytest.approx(np.eye(3), abs=1e-12)
        assert (dp.T @ (w[:, None] * dp)) == pytest.approx(np.eye(3) * H * H / 4, rel=1e-12, abs=1e-24)
        for n in range(27):
            node = np.array([idx[n] % g.nx, (idx[n] // g.nx) % g.ny, idx[n] // (g.nx * g.ny)]) * H
            assert dp[n] == pytest.approx(node - p, abs=1e-18)
            ref = bspline((p[0] - node[0]) / H) * bspline((p[1] - node[1]) / H) * bspline((p[2] - node[2]) / H)
            assert w[n] == pytest.approx(ref, rel=1e-12, abs=1e-16)


def one_particle(status, v, C9, stress=(0, 0, 0, 0, 0, 0)):
    P = workloads.params_dict(fixture=0.0)
    o = oracle.Oracle(**{k: v_ for k, v_ in P.items() if k in {f for f, _ in oracle.Params._fields_}})
    d = oracle.make_arrays(1)
    d["pos"][0] = [1.23e-3, 2.31e-3, 3.12e-3]
    d["vel"][0] = v
    d["mass"][0] = 2.0e-8
    d["volume"][0] = 4.0e-12
    d["density"][0] = 4430.0
    d["temperature"][0] = 400.0
    d["stress"][0] = stress
    d["status"][0] = status
    o.set_particles(d)
    o.set_transfer(oracle.TRANSFER_APIC)
    o.set_affine(np.asarray(C9, float).reshape(1, 9))
    o.step(1e-12)
    return o, d


def test_single_particle_p2g_against_textbook_bspline():
    C9 = np.array([30.0, -5.0, 2.0, 7.0, 11.0, -3.0, 1.0, 4.0, -9.0])
    v = np.array([0.5, -1.25, 2.0])
    o, d = one_particle(oracle.ACTIVE, v, C9)
    g = o.grid()
    snap = o.grid_snapshot()
    p = d["pos"][0]
    m = d["mass"][0]
    org = np.array([g.minx, g.miny, g.minz])
    r = (p - org) / H
    base = np.floor(r - 0.5).astype(int)
    total_m = 0.0
    Cm = C9.reshape(3, 3)
    for i in range(3):
        for j in range(3):
            for k in range(3):
                n = np.array([base[0] + i, base[1] + j, base[2] + k])
                xi = org + n * H
                w = np.prod([bspline(r[a] - n[a]) for a in range(3)])
                idx = (n[2] * g.ny + n[1]) * g.nx + n[0]
                assert snap["mass"][idx] == pytest.approx(m * w, rel=1e-12)
                assert snap["mom"][idx] == pytest.approx(m * w * (v + Cm @ (xi - p)), rel=1e-10, abs=1e-25)
                assert snap["tm"][idx] == pytest.approx(400.0 * m * w, rel=1e-12)
                total_m += m * w
    assert snap["mass"].sum() == pytest.approx(m, rel=1e-13)
    assert total_m == pytest.approx(m, rel=1e-13)
    assert np.count_nonzero(snap["mass"]) == 27


def test_fixed_particle_adds_mass_not_momentum():
    o, d = one_particle(oracle.FIXED_BOUNDARY, [1.0, 2.0, 3.0], np.ones(9) * 50.0)
    snap = o.grid_snapshot()
    assert snap["mass"].sum() == pytest.approx(d["mass"][0], rel=1e-13)
    assert np.abs(snap["mom"]).max() == 0.0


@pytest.mark.parametrize("transfer", [oracle.TRANSFER_APIC, oracle.TRANSFER_APIC_TEXTBOOK, oracle.TRANSFER_PIC])
def test_affine_field_step(transfer):
    cfg = workloads.scaled(workloads.slab_config0(), (6, 6, 4), steps=1)
    P = dict(cfg["params"], fixture_thickness=0.0)
    o = oracle.Oracle(**{k: v for k, v in P.items() if k in {f for f, _ in oracle.Params._fields_}})
    d = oracle.gen_jittered_slab(cfg["seed"], 6, 6, 4, H, (0, 0, 0), P["density"], 0.0)
    A = np.array([[40.0, -25.0, 10.0], [15.0, -30.0, 5.0], [-20.0, 35.0, -10.0]])
    v0 = np.array([0.3, -0.2, 0.1])
    x0 = d["pos"].copy()
    d["vel"][:] = v0 + x0 @ A.T
    d["temperature"][:] = 293.0
    o.set_particles(d)
    o.set_transfer(transfer)
    o.set_affine(np.tile(A.reshape(9), (len(x0), 1)))
    dt = 1e-9
    o.step(dt)
    out = o.get_particles()
    C9 = o.get_affine().reshape(-1, 3, 3)
    if transfer != oracle.TRANSFER_PIC:
        factor = 0.7425 if transfer == oracle.TRANSFER_APIC else 1.0
        assert out["vel"] == pytest.approx(v0 + x0 @ A.T, rel=1e-11, abs=1e-12)
        assert np.abs(C9 - factor * A).max() < 1e-10 * np.abs(A).max()
        assert out["pos"] == pytest.approx(x0 + dt * (v0 + x0 @ A.T), rel=0, abs=1e-18)
    else:
        # PIC: the affine matrices are not used or updated, and the trilinear
        # transfer does not reproduce the field at the slab's faces
        assert np.abs(C9 - A).max() == 0.0
        assert np.abs(out["vel"] - (v0 + x0 @ A.T)).max() > 1e-3
