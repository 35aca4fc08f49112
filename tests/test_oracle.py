"""Pins for the CPU oracle (oracle/mpm_oracle.c), independent of its own code.

Each test checks an oracle function against a closed form, an invariant, an
independent formula from elsewhere in the reference (MaterialModels.h), or a
hand-computed value -- chosen so that a dropped term, a wrong sign, a wrong
index or a transposed operand fails.  The reference ships no tests, golden
vectors or fixtures for this path.
"""
import math

import numpy as np
import pytest


def _grid(orc, cell=2.5e-4, mn=(-1e-3, -1e-3, -1e-3), n=(20, 18, 16)):
    g = orc.Grid()
    g.cell = cell
    g.minx, g.miny, g.minz = mn
    g.nx, g.ny, g.nz = n
    return g


def test_weights_partition_linear_reproduction(orc):
    g = _grid(orc)
    rng = np.random.default_rng(1)
    for _ in range(50):
        x = -1e-3 + rng.uniform(0.5, 14.5, size=3) * g.cell
        idx, w, gr = orc.weights(g, *x)
        assert (idx >= 0).all()
        i, j, k = idx % g.nx, (idx // g.nx) % g.ny, idx // (g.nx * g.ny)
        xn = np.stack([g.minx + i * g.cell, g.miny + j * g.cell, g.minz + k * g.cell], 1)
        assert abs(w.sum() - 1.0) < 1e-14
        assert np.abs(gr.sum(0)).max() < 1e-9            # sum of gradients is zero
        assert np.allclose((w[:, None] * xn).sum(0), x, rtol=0, atol=1e-18)
        # gradient of the linear field x_a reproduced exactly: sum x_i (x) grad w = I
        assert np.allclose(xn.T @ gr, np.eye(3), atol=1e-12)
        # node order: dz outer, dy, dx inner
        assert idx[1] - idx[0] == 1 and idx[2] - idx[0] == g.nx and idx[4] - idx[0] == g.nx * g.ny


def test_weights_out_of_grid_nodes(orc):
    g = _grid(orc)
    x = np.array([g.minx + (g.nx - 1 + 0.3) * g.cell, g.miny + 2.5 * g.cell, g.minz + 2.5 * g.cell])
    idx, w, gr = orc.weights(g, *x)
    assert (idx[1::2] == -1).all() and (w[1::2] == 0).all() and (gr[1::2] == 0).all()
    assert (idx[0::2] >= 0).all()


def test_equivalent_stress_closed_forms(orc):
    assert orc.equivalent_stress(5e8, 0, 0, 0, 0, 0) == pytest.approx(5e8, rel=1e-15)
    assert orc.equivalent_stress(0, 0, 0, 2e8, 0, 0) == pytest.approx(math.sqrt(3) * 2e8, rel=1e-15)
    assert orc.equivalent_stress(7e8, 7e8, 7e8, 0, 0, 0) == 0.0
    rng = np.random.default_rng(2)
    for _ in range(20):
        s = rng.normal(size=6) * 1e8
        S = np.array([[s[0], s[3], s[4]], [s[3], s[1], s[5]], [s[4], s[5], s[2]]])
        dev = S - np.trace(S) / 3 * np.eye(3)
        assert orc.equivalent_stress(*s) == pytest.approx(math.sqrt(1.5 * (dev * dev).sum()), rel=1e-12)


def _jc_materialmodels(A, B, n, C, m, rate_ref, Tr, Tm, strain, rate, T):
    # calculateJohnsonCookStress, include/MaterialModels.h:71-94 (independent restatement)
    st = A + B * abs(strain) ** n
    r = rate / rate_ref
    if r < 1.0:
        r = 1.0
    rt = 1.0 + C * math.log(r)
    ts = 0.0
    if T > Tr:
        ts = (T - Tr) / (Tm - Tr)
        ts = min(ts, 1.0)
    th = 1.0 - ts ** m
    th = max(th, 0.01)
    return st * rt * th


def test_jc_flow_against_materialmodels_and_closed_forms(orc):
    p = orc.default_params(jc_A=1098e6, jc_B=1092e6, jc_n=0.93, jc_C=0.014, jc_m=1.1,
                           ref_temp=293.0, melt_temp=1878.0, ref_strain_rate=1.0)
    assert orc.jc_flow(p, 0.0, 0.5, 293.0) == pytest.approx(1098e6, rel=1e-15)
    assert orc.jc_flow(p, 1.0, 1.0, 293.0) == pytest.approx(1098e6 + 1092e6, rel=1e-15)
    assert orc.jc_flow(p, 0.0, math.e, 293.0) == pytest.approx(1098e6 * 1.014, rel=1e-14)
    assert orc.jc_flow(p, 0.0, 1.0, 5000.0) == pytest.approx(1098e6 * 0.01, rel=1e-14)
    rng = np.random.default_rng(3)
    for _ in range(50):
        e = rng.uniform(0, 2)
        r = 10 ** rng.uniform(-2, 6)
        T = rng.uniform(293, 1900)
        ref = _jc_materialmodels(1098e6, 1092e6, 0.93, 0.014, 1.1, 1.0, 293.0, 1878.0, e, r, T)
        assert orc.jc_flow(p, e, r, T) == pytest.approx(ref, rel=1e-13)


def test_rebuild_params_hand_computed(orc):
    g = orc.rebuild_params(0.25, (0.0, 0.0, 0.0), (1.1, 2.0, 0.3))
    assert (g.minx, g.miny, g.minz) == (-1.0, -1.0, -1.0)
    assert (g.nx, g.ny, g.nz) == (14, 17, 11)      # ceil(12.4)+1, ceil(16)+1, ceil(9.2)+1


def test_chip_kick_u32_hash(orc):
    # 1 * 2654435761 = 0x9E3779B1 -> bytes 0x37, 0x79, 0xB1
    assert np.allclose(orc.chip_kick(1), [55 / 255 - 0.5, 121 / 255 - 0.5, 177 / 255 - 0.5], atol=0)
    # (uint32)(-1 * 2654435761) = 0x61C8864F -> bytes 0xC8, 0x86, 0x4F
    assert np.allclose(orc.chip_kick(-1), [200 / 255 - 0.5, 134 / 255 - 0.5, 79 / 255 - 0.5], atol=0)


def _lame(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    G = E / (2 * (1 + nu))
    return lam, G


def _particle(orc, **kw):
    p = orc.Particle()
    p.mass, p.volume, p.temperature = 1e-9, 2e-12, 293.0
    for k, v in kw.items():
        if k == "s":
            for i, x in enumerate(v):
                p.s[i] = x
        else:
            setattr(p, k, v)
    return p


def test_constitutive_uniaxial_and_shear_hooke(orc):
    prm = orc.default_params(young=113.8e9, poisson=0.342, ref_temp=293.0, melt_temp=1878.0)
    lam, G = _lame(113.8e9, 0.342)
    dt, e = 1e-8, 100.0
    p = orc.constitutive(prm, _particle(orc), np.diag([e, 0, 0]), dt)
    assert p.s[0] == pytest.approx((lam + 2 * G) * e * dt, rel=1e-12)
    assert p.s[1] == pytest.approx(lam * e * dt, rel=1e-12)
    assert p.s[2] == pytest.approx(lam * e * dt, rel=1e-12)
    assert p.strainRate == pytest.approx(2 * e / 3, rel=1e-12)
    assert p.damage == pytest.approx(2 * e / 3 * dt / 0.3, rel=1e-12)
    assert p.plasticStrain == 0.0
    gam = 50.0
    L = np.zeros((3, 3)); L[0, 1] = gam
    p = orc.constitutive(prm, _particle(orc), L, dt)
    assert p.s[3] == pytest.approx(G * gam * dt, rel=1e-12)
    assert p.s[0] == 0 and p.s[4] == 0 and p.s[5] == 0
    assert p.strainRate == pytest.approx(gam / math.sqrt(3), rel=1e-12)
    # rigid spin: no stress, no rate, no damage
    L = np.zeros((3, 3)); L[0, 1], L[1, 0] = 3e3, -3e3
    p = orc.constitutive(prm, _particle(orc, s=[1e8, 0, 0, 0, 0, 0]), L, dt)
    assert p.s[0] == 1e8 and p.strainRate == 0.0 and p.damage == 0.0


def test_constitutive_radial_return(orc):
    prm = orc.default_params(young=113.8e9, poisson=0.342, ref_temp=293.0, melt_temp=1878.0,
                             jc_A=1098e6, jc_B=1092e6, jc_n=0.93, jc_C=0.014, jc_m=1.1,
                             density=4430.0, specific_heat=526.0, max_contact_temp_rise=50.0)
    lam, G = _lame(113.8e9, 0.342)
    dt, e = 1e-8, 4e3
    s0 = np.array([1.2e9, 0, 0, 0, 0, 0])
    p = orc.constitutive(prm, _particle(orc, s=list(s0), temperature=400.0, plasticStrain=0.01),
                         np.diag([e, 0, 0]), dt)
    trial = s0 + np.array([(lam + 2 * G) * e * dt, lam * e * dt, lam * e * dt, 0, 0, 0])
    q_trial = orc.equivalent_stress(*trial)
    flow = _jc_materialmodels(1098e6, 1092e6, 0.93, 0.014, 1.1, 1.0, 293.0, 1878.0, 0.01, 2 * e / 3, 400.0)
    assert q_trial > flow
    sn = np.array(p.s[:])
    assert orc.equivalent_stress(*sn) == pytest.approx(flow, rel=1e-12)     # back on the surface
    assert np.allclose(sn, trial * flow / q_trial, rtol=1e-12)             # scales the pressure too
    dp = (q_trial - flow) / (3 * G)
    assert p.plasticStrain == pytest.approx(0.01 + dp, rel=1e-12)
    assert p.temperature == pytest.approx(400.0 + min(0.9 * flow * dp / (4430 * 526.0), 50.0), rel=1e-12)
    # kinematic hardening moves the backstress along the shifted stress
    prm.kinematic_hardening = 2e9
    p = orc.constitutive(prm, _particle(orc, s=list(s0), temperature=400.0, plasticStrain=0.01),
                         np.diag([e, 0, 0]), dt)
    assert np.allclose(np.array(p.a[:]), 2e9 * dp / q_trial * trial, rtol=1e-10)


def test_constitutive_cdm_and_chip_transition(orc):
    prm = orc.default_params(ref_temp=293.0, melt_temp=1878.0, damage_threshold=1.0,
                             critical_plastic_strain=0.15, thermal_chip_ratio=0.25)
    s = [3e8, -1e8, 0, 5e7, 0, 0]
    p = orc.constitutive(prm, _particle(orc, s=s, damage=0.8, residualStress=1.0), np.zeros((3, 3)), 1e-8)
    g = (1.0 - 0.8) / 0.3
    assert np.allclose(p.s[:], np.array(s) * g, rtol=1e-14)
    assert p.residualStress == pytest.approx(orc.equivalent_stress(*(np.array(s) * g)), rel=1e-14)
    assert p.status == orc.ACTIVE
    p = orc.constitutive(prm, _particle(orc, s=s, plasticStrain=0.2), np.zeros((3, 3)), 1e-8)
    assert p.status == orc.CHIP and list(p.s[:]) == [0.0] * 6 and p.damage == 1.0
    assert p.residualStress == pytest.approx(orc.equivalent_stress(*s), rel=1e-14)
    p = orc.constitutive(prm, _particle(orc, temperature=0.25 * 1878.0), np.zeros((3, 3)), 1e-8)
    assert p.status == orc.CHIP
    p = orc.constitutive(prm, _particle(orc, temperature=0.25 * 1878.0 - 1e-6), np.zeros((3, 3)), 1e-8)
    assert p.status == orc.ACTIVE


def _slab(orc, cx=6, cy=5, cz=4, h=2.5e-4, fixture=0.0, seed=7):
    return orc.gen_jittered_slab(seed, cx, cy, cz, h, (0.0, 0.0, 0.0), 4430.0, fixture)


def _on_node_plane(g, pos):
    r = (pos - np.array([g.minx, g.miny, g.minz])) / g.cell
    return np.any(r == np.floor(r), axis=1)


def test_node_plane_quirk_spurious_gradient(orc):
    # A particle exactly on a node keeps only that node (w=1); the skipped
    # nodes' gradients vanish from L = sum v_n (x) grad w_n, so a rigid
    # translation v0*e_x gives L[0,:] = -v0/h*(1,1,1) and
    # eqRate = sqrt(2/3*(4/9+1/9+1/9+2*(1/4+1/4)))*v0/h = sqrt(10)/3*v0/h.
    h, v0, dt = 0.25, 2.0, 1e-6
    d = orc.make_arrays(1)
    d["vel"][0] = [v0, 0, 0]
    d["mass"][0], d["volume"][0], d["temperature"][0] = 1.0, 1e-3, 300.0
    s = orc.Oracle(cell_size=h, density=1000.0, specific_heat=500.0, thermal_conductivity=0.0,
                   ref_temp=0.0, melt_temp=1e4, fixture_thickness=-100.0, young=1e6, poisson=0.25)
    s.set_particles(d)
    s.step(dt)
    o = s.get_particles()
    assert o["strain_rate"][0] == pytest.approx(math.sqrt(10) / 3 * v0 / h, rel=1e-14)
    lam, G = _lame(1e6, 0.25)
    Dxx, Dxy = -v0 / h, -v0 / (2 * h)
    assert o["stress"][0][0] == pytest.approx((lam + 2 * G) * Dxx * dt, rel=1e-12)
    assert o["stress"][0][1] == pytest.approx(lam * Dxx * dt, rel=1e-12)
    assert o["stress"][0][3] == pytest.approx(2 * G * Dxy * dt, rel=1e-12)
    assert o["stress"][0][5] == 0.0


def _solver(orc, **kw):
    base = dict(cell_size=2.5e-4, fixture_thickness=0.0, ref_temp=293.0, melt_temp=1878.0,
                density=4430.0, specific_heat=526.0, young=113.8e9, poisson=0.342,
                max_particles=10 ** 8)
    base.update(kw)
    return orc.Oracle(**base)


def test_p2g_conservation(orc):
    d = _slab(orc)
    rng = np.random.default_rng(4)
    n = len(d["mass"])
    d["vel"][:] = rng.normal(size=(n, 3))
    d["temperature"][:] = rng.uniform(293, 400, n)
    d["stress"][:] = rng.normal(size=(n, 6)) * 1e8
    d["fext"][:] = rng.normal(size=(n, 3)) * 1e-3
    s = _solver(orc)
    s.set_particles(d)
    # particles exactly on a node plane skip their w=0 nodes (and those nodes'
    # non-zero gradients, MPMWorkpieceSolver.cu:112): give them no stress here
    d["stress"][_on_node_plane(s.grid(), d["pos"])] = 0.0
    s.set_particles(d)
    s.step(1e-9)
    g = s.grid_snapshot()
    m = d["mass"]
    assert g["mass"].sum() == pytest.approx(m.sum(), rel=1e-13)
    assert np.allclose(g["mom"].sum(0), (m[:, None] * d["vel"]).sum(0), rtol=1e-12, atol=1e-22)
    assert g["tm"].sum() == pytest.approx((m * d["temperature"]).sum(), rel=1e-13)
    # stress divergence sums to zero per particle; only f_ext survives
    assert np.allclose(g["force"].sum(0), d["fext"].sum(0), rtol=1e-9, atol=1e-12)


def test_rigid_translation(orc):
    d = _slab(orc)
    d["vel"][:] = [-1.0, 0.25, 0.1]
    d["temperature"][:] = 293.0
    s = _solver(orc)
    s.set_particles(d)
    dt = 1e-8
    s.step(dt)
    o = s.get_particles()
    plane = _on_node_plane(s.grid(), d["pos"])
    assert 0 < plane.sum() < len(plane)
    assert np.allclose(o["vel"], [-1.0, 0.25, 0.1], rtol=1e-13, atol=1e-15)
    assert np.allclose(o["pos"], d["pos"] + dt * np.array([-1.0, 0.25, 0.1]), rtol=0, atol=1e-19)
    assert np.abs(o["stress"][~plane]).max() < 1e-3
    assert np.abs(o["stress"][plane]).max() > 1e5      # the node-plane quirk, pinned below
    # surface cooling toward empty nodes is undone by the clamp T >= T_ref
    assert np.allclose(o["temperature"], 293.0, rtol=1e-14, atol=0)


def test_fixture_nodes_and_fixed_particles(orc):
    d = _slab(orc, fixture=5e-4)
    d["vel"][:] = [0.0, 0.0, 1.0]
    s = _solver(orc, fixture_thickness=5e-4)
    s.set_particles(d)
    s.step(1e-8)
    o = s.get_particles()
    fixed = d["status"] == orc.FIXED_BOUNDARY
    assert fixed.any() and (~fixed).any()
    assert np.all(o["pos"][fixed] == d["pos"][fixed]) and np.all(o["vel"][fixed] == 0)
    st = s.grid_state()
    g = s.grid()
    k = np.arange(len(st["mass"])) // (g.nx * g.ny)
    z = g.minz + k * g.cell
    live = st["mass"] > 1e-18
    expect_fixed = z <= g.minz + 5e-4 + g.cell
    assert np.all(st["fixed"][live] == expect_fixed[live])
    assert np.all(st["vel"][live & expect_fixed] == 0)


def test_diffusion_closed_forms(orc):
    # one particle exactly on a node (6 empty neighbours) and one at a cell
    # centre (each of its 8 nodes has 3 empty neighbours); alpha*dt/h^2 = 0.01
    h, dt, rho, cp = 0.25, 1e-3, 1000.0, 500.0
    k = 0.01 * rho * cp * h * h / dt
    d = orc.make_arrays(2)
    d["pos"][:] = [[-2.0, -2.0, -2.0], [0.125, 0.125, 0.125]]
    d["mass"][:] = 1.0
    d["volume"][:] = 1e-3
    d["temperature"][:] = [400.0, 300.0]
    d["id"][:] = [0, 1]
    s = orc.Oracle(cell_size=h, density=rho, specific_heat=cp, thermal_conductivity=k,
                   ref_temp=0.0, melt_temp=1e4, fixture_thickness=-100.0)
    s.set_particles(d)
    s.step(dt)
    o = s.get_particles()
    assert o["temperature"][0] == pytest.approx(400.0 * 0.94, rel=1e-13)
    assert o["temperature"][1] == pytest.approx(300.0 * 0.97, rel=1e-13)


def _one_particle_at_node(orc, v, T=300.0):
    d = orc.make_arrays(1)
    d["pos"][0] = [0.0, 0.0, 0.0]
    d["vel"][0] = v
    d["mass"][0], d["volume"][0], d["temperature"][0] = 2.0, 1e-3, T
    return d


def test_point_tool_contact_projection(orc):
    h, dt = 0.25, 1e-3
    n0 = np.array([1.0, 2.0, 2.0]) / 3.0
    t0 = np.array([2.0, -2.0, 1.0]) / 3.0          # orthogonal to n0
    v = -1.5 * n0 + 0.7 * t0
    d = _one_particle_at_node(orc, v)
    s = orc.Oracle(cell_size=h, density=1000.0, specific_heat=500.0, thermal_conductivity=0.0,
                   ref_temp=0.0, melt_temp=1e4, fixture_thickness=-100.0, contact_damping=0.15,
                   max_contact_temp_rise=50.0)
    s.set_particles(d)
    samples = np.array([-0.2 * n0,                  # nearest: hit, normal n0
                        -0.3 * t0,                   # farther: ignored
                        -5.0 * n0])                  # outside r_c
    s.set_tool_samples(samples, 0.3)
    s.step(dt)
    o = s.get_particles()
    assert np.allclose(o["vel"][0], 0.7 * t0, atol=1e-14)
    assert s.metrics().contact_force == pytest.approx(2.0 * 1.5 / dt, rel=1e-14)
    dT = min(0.5 * 2.0 * 1.5 ** 2 * 0.15 / (2.0 * 500.0), 50.0)
    assert o["temperature"][0] == pytest.approx(300.0 + dT, rel=1e-14)
    # moving away from the sample: untouched
    d = _one_particle_at_node(orc, 1.5 * n0)
    s.set_particles(d)
    s.step(dt)
    assert np.allclose(s.get_particles()["vel"][0], 1.5 * n0, atol=1e-14)


def test_point_tool_stride_rule(orc):
    # 8001 mesh nodes -> stride 2: only even-indexed nodes are samples
    h, dt = 0.25, 1e-3
    n0 = np.array([0.0, 0.0, 1.0])
    d = _one_particle_at_node(orc, -1.0 * n0)
    s = orc.Oracle(cell_size=h, density=1000.0, specific_heat=500.0, thermal_conductivity=0.0,
                   ref_temp=0.0, melt_temp=1e4, fixture_thickness=-100.0)
    s.set_particles(d)
    pts = np.tile(np.array([[10.0, 10.0, 10.0]]), (8001, 1))
    pts[1::2] = -0.1 * n0                            # odd nodes would hit
    s.set_tool_samples(pts, 0.3)
    s.step(dt)
    assert np.allclose(s.get_particles()["vel"][0], -n0, atol=1e-14)


def test_chip_separation_and_fracture(orc):
    h = 0.25
    d = orc.make_arrays(3)
    d["pos"][:] = [[0.0, 0.0, 0.0], [0.0, 0.0, 2.0], [3.0, 0.0, 0.0]]
    d["mass"][:] = 1.0
    d["volume"][:] = 1e-3
    d["temperature"][:] = 300.0
    d["plastic_strain"][:] = [0.01, 0.01, 0.0]
    d["residual_stress"][:] = [7.0, 0.0, 0.0]
    d["damage"][:] = [0.2, 0.2, 1.6]
    d["status"][:] = [orc.ACTIVE, orc.ACTIVE, orc.CHIP]
    d["id"][:] = [5, 6, 1]
    s = orc.Oracle(cell_size=h, density=1000.0, specific_heat=500.0, thermal_conductivity=0.0,
                   ref_temp=0.0, melt_temp=1e4, fixture_thickness=-100.0, damage_threshold=1.0,
                   critical_plastic_strain=0.15, thermal_chip_ratio=0.9)
    s.set_particles(d)
    s.set_tool_tip(0.0, 0.0, 1.0)                   # cut depth 1.5h = 0.375 -> z < 0.625
    s.step(1e-6)                                     # step index 0: even
    o = s.get_particles()
    assert o["status"][0] == orc.CHIP and o["damage"][0] == 1.0 and o["residual_stress"][0] == 7.0
    assert o["status"][1] == orc.ACTIVE
    assert np.allclose(o["vel"][2], 0.5 * np.array([55 / 255 - 0.5, 121 / 255 - 0.5, 177 / 255 - 0.5]))
    assert o["damage"][2] == 0.5
    s.step(1e-6)                                     # odd step: no fracture pass
    assert s.get_particles()["damage"][2] == 0.5


def test_stable_dt(orc):
    s = orc.Oracle(cell_size=2.5e-4, young=113.8e9, density=4430.0)
    assert s.stable_dt() == pytest.approx(0.35 * 2.5e-4 / math.sqrt(113.8e9 / 4430.0), rel=1e-15)
    s = orc.Oracle(cell_size=2.5e-4, young=113.8e9, density=4430.0, min_time_step=1e-7)
    assert s.stable_dt() == 1e-7


def test_box_init_lattice_and_fixture(orc):
    s = orc.Oracle(cell_size=1e-3, fixture_thickness=1.5e-3, max_particles=10 ** 6, ref_temp=20.0)
    s.init_box((0, 0, 0), (4e-3, 3e-3, 5e-3), 1e-3)
    o = s.get_particles()
    assert len(o["mass"]) == 5 * 4 * 6
    assert o["pos"][0].tolist() == [0.0, 0.0, 5e-3]         # z descends from the top, x fastest
    assert o["pos"][1][0] == pytest.approx(1e-3)
    assert np.all(o["status"][o["pos"][:, 2] <= 1.5e-3 + 1e-12] == orc.FIXED_BOUNDARY)
    assert np.all(o["temperature"] == 20.0) and np.all(o["id"] == np.arange(120))
    assert o["mass"][0] == pytest.approx(4430.0 * 1e-9)


def test_metrics_every_25_steps_and_compaction(orc):
    d = _slab(orc, cx=3, cy=3, cz=2)
    d["vel"][:] = [0.5, 0.0, 0.0]
    d["temperature"][:] = 293.0
    d["status"][::7] = orc.INACTIVE
    s = _solver(orc)
    s.set_particles(d)
    for _ in range(24):
        s.step(1e-8)
    assert s.metrics().kinetic_energy == 0.0
    s.step(1e-8)
    o = s.get_particles()
    live = o["status"] != orc.INACTIVE
    ke = 0.5 * (o["mass"][live] * (o["vel"][live] ** 2).sum(1)).sum()
    assert s.metrics().kinetic_energy == pytest.approx(ke, rel=1e-12)
    n0 = s.count()
    s.compact()
    o2 = s.get_particles()
    assert s.count() == live.sum() < n0
    assert np.all(o2["id"] == o["id"][live])


def test_wedge_and_drill_sdf(orc):
    desc = orc.ToolDesc()
    desc.kind = orc.TOOL_WEDGE
    desc.rake_angle, desc.clearance_angle = math.radians(5), math.radians(7)
    pose = orc.ToolPose()
    pose.center[:] = [1.0, 2.0, 3.0]
    g = math.radians(5)
    nr = np.array([math.cos(g), 0, math.sin(g)])
    for dd in (1e-3, -2e-4):
        x = np.array([1.0, 2.0, 3.0]) + dd * nr + np.array([0, 0.3, 0]) + 5e-3 * np.array([-math.sin(g), 0, math.cos(g)])
        assert orc.tool_phi(desc, pose, *x) == pytest.approx(dd, abs=1e-14)
    desc.grad_eps = 1e-6
    x = np.array([1.0, 2.0, 3.0]) + 5e-3 * np.array([-math.sin(g), 0, math.cos(g)])
    assert np.allclose(orc.tool_normal(desc, pose, *x), nr, atol=1e-8)
    pose.angle = math.pi / 2                      # rake normal now along +y
    x = np.array([1.0, 2.0, 3.0]) + 1e-3 * np.array([0, math.cos(g), math.sin(g)]) + 5e-3 * np.array([0, -math.sin(g), math.cos(g)])
    assert orc.tool_phi(desc, pose, *x) == pytest.approx(1e-3, abs=1e-12)

    dr = orc.ToolDesc()
    dr.kind = orc.TOOL_DRILL
    dr.radius, dr.helix_angle, dr.point_angle = 5e-3, math.radians(30), math.radians(118)
    dr.flute_offset, dr.flute_radius, dr.flute_length = 3e-3, 2.25e-3, 20e-3
    p0 = orc.ToolPose()
    assert orc.tool_phi(dr, p0, 0, 0, -2e-3) == pytest.approx(2e-3, rel=1e-14)     # below the tip
    # a point on the cylinder, 90 degrees from the flute plane at this height
    z = 10e-3
    psi = z * math.tan(math.radians(30)) / 5e-3 + math.pi / 2
    assert orc.tool_phi(dr, p0, 5e-3 * math.cos(psi), 5e-3 * math.sin(psi), z) == pytest.approx(0.0, abs=1e-15)
    rng = np.random.default_rng(5)
    for _ in range(20):
        q = rng.uniform(-6e-3, 6e-3, 3); q[2] = rng.uniform(4e-3, 15e-3)
        a = orc.tool_phi(dr, p0, *q)
        assert orc.tool_phi(dr, p0, -q[0], -q[1], q[2]) == pytest.approx(a, abs=1e-15)  # two flutes
        dz = 1e-3                                                  # helical invariance
        th = dz * math.tan(math.radians(30)) / 5e-3
        qr = [q[0] * math.cos(th) - q[1] * math.sin(th), q[0] * math.sin(th) + q[1] * math.cos(th), q[2] + dz]
        assert orc.tool_phi(dr, p0, *qr) == pytest.approx(a, abs=1e-12)


def test_analytic_tool_contact_relative_velocity_and_friction(orc):
    h, dt = 0.25, 1e-3
    desc = orc.ToolDesc()
    desc.kind, desc.band, desc.friction, desc.grad_eps = orc.TOOL_WEDGE, 0.1, 0.5, 1e-4
    desc.rake_angle, desc.clearance_angle = 0.0, 0.0
    pose = orc.ToolPose()
    pose.center[:] = [-0.05, 0.0, -1.0]           # rake plane x = -0.05, node at x=0: phi = 0.05 < band
    pose.vel[:] = [0.2, 0.0, 0.0]
    d = _one_particle_at_node(orc, [-1.0, 0.4, 0.0])
    s = orc.Oracle(cell_size=h, density=1000.0, specific_heat=500.0, thermal_conductivity=0.0,
                   ref_temp=0.0, melt_temp=1e4, fixture_thickness=-100.0)
    s.set_particles(d)
    s.set_tool_analytic(desc)
    s.set_tool_pose(pose)
    s.step(dt)
    v = s.get_particles()["vel"][0]
    # relative normal velocity -1.2 removed; tangential 0.4 reduced by mu*1.2 = 0.6 -> 0
    assert np.allclose(v, [0.2, 0.0, 0.0], atol=1e-9)
    desc.friction = 0.1
    s.set_tool_analytic(desc)
    pose.center[:] = [-0.05, 0.0, -1.0]
    s.set_tool_pose(pose)
    s.set_particles(_one_particle_at_node(orc, [-1.0, 0.4, 0.0]))
    s.step(dt)
    v = s.get_particles()["vel"][0]
    assert np.allclose(v, [0.2, 0.4 - 0.12, 0.0], atol=1e-9)


def test_jittered_slab_generator(orc):
    d = orc.gen_jittered_slab(0xED6E0001, 4, 3, 2, 2.5e-4, (0.0, 0.0, 0.0), 4430.0, 1e-4)
    n = 4 * 3 * 2 * 8
    assert len(d["mass"]) == n and np.all(d["id"] == np.arange(n))
    pos = d["pos"]
    assert np.all(pos.astype(np.float32).astype(np.float64) == pos)
    c = np.arange(n) // 8
    k = np.arange(n) % 8
    cell = np.stack([c % 4, (c // 4) % 3, c // 12], 1)
    sub = np.stack([k & 1, (k >> 1) & 1, (k >> 2) & 1], 1)
    lo = (cell + 0.5 * sub) * 2.5e-4
    assert np.all(pos >= lo - 1e-12) and np.all(pos <= lo + 1.25e-4 + 1e-12)
    assert np.allclose(d["volume"], 2.5e-4 ** 3 / 8)
    assert np.all((d["status"] == 4) == (pos[:, 2] <= 1e-4))
    d2 = orc.gen_jittered_slab(0xED6E0001, 4, 3, 2, 2.5e-4, (0.0, 0.0, 0.0), 4430.0, 1e-4)
    assert np.array_equal(d2["pos"], pos)
