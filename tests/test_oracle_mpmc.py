"""Pins of the tool FEM and hash-contact oracle (oracle/mpmc_oracle.c, the
restatement of FEMSolver::step, src/FEMSolver.cu, and of ContactSolver's
hash path, src/ContactSolver.cu / SharedSpatialHash.cu) by closed forms,
invariants and hand-derived cases, on the CPU.  The known answers are
derived here from the mathematics, not retyped from the oracle."""
import math

import numpy as np
import pytest

import mpmc
import mpms
import toolfem

H = 5e-4


def single_tet(a=1e-3, flip=False):
    """The right tet (0,0,0), (a,0,0), (0,a,0), (0,0,a) as tet and surface
    nodes (each surface node embedded in its tet node with weight 1)."""
    xyz = np.array([[0, 0, 0], [a, 0, 0], [0, a, 0], [0, 0, a]], float)
    el = np.array([[0, 2, 1, 3]] if flip else [[0, 1, 2, 3]], np.int32)
    emb_tet = np.full((4, 4), -1, np.int32)
    emb_tet[:, 0] = np.arange(4)
    emb_w = np.zeros((4, 4))
    emb_w[:, 0] = 1.0
    return {"surf_xyz": xyz, "surf_region": np.zeros(4, np.int32), "tet_xyz": xyz,
            "tet_region": np.ones(4, np.int32), "elems": el, "emb_surf": np.arange(4, dtype=np.int32),
            "emb_tet": emb_tet, "emb_w": emb_w}


def test_element_setup_volume_inverse_shape_masses():
    """createElementsFromMesh (:1012-1078): V = |det Dm| / 6 = a^3 / 6, a
    negatively oriented tet is reordered, invDm Dm = I, each node gets
    rho V / 4; surface masses from the embeds (:973-992); tets within 5 mm of
    the top are fixed (:955-962)."""
    a = 1e-3
    for flip in (False, True):
        T = mpmc.ToolOracle(density=8000.0)
        T.set_mesh(single_tet(a, flip))
        n4, vol, inv = T.elements()
        assert vol[0] == pytest.approx(a ** 3 / 6, rel=1e-14)
        X = T.tets()
        P = np.stack([X["ox"], X["oy"], X["oz"]], 1)[n4[0]]
        Dm = np.stack([P[1] - P[0], P[2] - P[0], P[3] - P[0]], 1)
        assert np.linalg.det(Dm) > 0
        assert np.allclose(inv[0] @ Dm, np.eye(3), atol=1e-12)
        assert np.allclose(X["mass"], 8000.0 * a ** 3 / 6 / 4, rtol=1e-14)
        assert np.allclose(T.surface()["mass"], X["mass"], rtol=1e-14)
        assert X["fixed"].all()            # a 1 mm tet lies within 5 mm of its top


def test_polar_textbook_and_reference_quirk():
    """The element polar iteration (:217-278).  Textbook (X^-T): R is a
    rotation and R^T F symmetric positive definite.  The reference adds X^-1
    instead: for a symmetric F the two agree; for a pure rotation Q by 30
    degrees about z it returns I, the rotation of sym(Q) = cos30 (I - e_z
    e_z^T) + e_z e_z^T, which is symmetric positive definite."""
    rng = np.random.default_rng(2)
    for _ in range(10):
        F = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
        if np.linalg.det(F) < 0.2:
            continue
        R = mpmc.polar(F, textbook=True)
        assert np.allclose(R.T @ R, np.eye(3), atol=1e-9)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-9)
        S = R.T @ F
        assert np.allclose(S, S.T, atol=1e-8) and np.all(np.linalg.eigvalsh(0.5 * (S + S.T)) > 0)
        Fs = 0.5 * (F + F.T) + np.eye(3)
        assert np.allclose(mpmc.polar(Fs), mpmc.polar(Fs, textbook=True), atol=1e-12)
    c, s = math.cos(math.pi / 6), math.sin(math.pi / 6)
    Q = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    assert np.allclose(mpmc.polar(Q, textbook=True), Q, atol=1e-12)
    assert np.allclose(mpmc.polar(Q), np.eye(3), atol=1e-9)


def _free_tet_tool(a=1e-3, **kw):
    """One tet, every node free (unfixed, not driven), no damping."""
    T = mpmc.ToolOracle(damping_ratio=0.0, **kw)
    T.set_mesh(single_tet(a))
    T.init_driven(-1.0)               # threshold above the top: none driven, all unfixed
    return T


def _place(T, X):
    t = T.tets()
    t["x"], t["y"], t["z"] = X[:, 0], X[:, 1], X[:, 2]
    T.set_tets(t)


def test_element_forces_uniaxial_stretch_closed_form():
    """computeElementForcesKernel (:167-342), normalizeNodeStressKernel
    (:346-369), integrateNodesKernel (:525-571): a tet stretched to
    X0 diag(1 + e, 1, 1) has F symmetric, R = I, Cauchy stress
    lambda e I + 2 mu e e_x e_x^T at every node (one element: the weights
    cancel), von Mises 0.7 of its equivalent, and nodal forces
    -V0 sigma invDm^T (node 0 the negative sum) giving v = f dt / m in one
    step; the forces sum to zero."""
    a, e, dt = 1e-3, 1e-4, 1e-8
    T = _free_tet_tool(a)
    X0 = single_tet(a)["tet_xyz"]
    _place(T, X0 * np.array([1 + e, 1, 1]))
    m0 = T.tets()["mass"]
    T.step(dt)
    t = T.tets()
    p = T.params
    mu = p.young / (2 * (1 + p.poisson))
    lam = p.young * p.poisson / ((1 + p.poisson) * (1 - 2 * p.poisson))
    sig = lam * e * np.eye(3) + 2 * mu * e * np.diag([1.0, 0.0, 0.0])
    voigt = np.array([sig[0, 0], sig[1, 1], sig[2, 2], sig[0, 1], sig[0, 2], sig[1, 2]])
    assert np.allclose(t["s"], voigt, rtol=1e-9, atol=1e-6 * abs(sig).max())
    vm = math.sqrt(0.5 * ((sig[0, 0] - sig[1, 1]) ** 2 + (sig[1, 1] - sig[2, 2]) ** 2 + (sig[2, 2] - sig[0, 0]) ** 2))
    assert np.allclose(t["von_mises"], 0.7 * vm, rtol=1e-9)
    invDm = np.linalg.inv((X0[1:] - X0[0]).T)
    Hm = -(a ** 3 / 6) * sig @ invDm.T
    f = np.vstack([-(Hm[:, 0] + Hm[:, 1] + Hm[:, 2]), Hm.T])
    v = np.stack([t["vx"], t["vy"], t["vz"]], 1)
    assert np.allclose(v, f / m0[:, None] * dt, rtol=1e-9, atol=1e-9 * np.abs(f / m0[:, None] * dt).max())
    assert np.allclose((v * m0[:, None]).sum(0), 0.0, atol=1e-12 * np.abs(v * m0[:, None]).max())


def test_element_forces_rigid_motions():
    """A rigid translation makes no force; a rigid rotation by 30 degrees
    about z makes none with the textbook polar (R = Q, eps = 0) and a
    non-zero one with the reference's (R = I there, eps = Q - I)."""
    a, dt = 1e-3, 1e-8
    X0 = single_tet(a)["tet_xyz"]
    T = _free_tet_tool(a)
    _place(T, X0 + np.array([3e-4, -2e-4, 1e-4]))
    T.step(dt)
    assert np.abs(np.stack([T.tets()[k] for k in ("vx", "vy", "vz")])).max() < 1e-9
    c, s = math.cos(math.pi / 6), math.sin(math.pi / 6)
    Q = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    for textbook, moves in ((1, False), (0, True)):
        T = _free_tet_tool(a, polar_textbook=textbook)
        _place(T, X0 @ Q.T)
        T.step(dt)
        vmax = np.abs(np.stack([T.tets()[k] for k in ("vx", "vy", "vz")])).max()
        assert (vmax > 1e-3) == moves, (textbook, vmax)


def test_spindle_spring_and_integration_closed_form():
    """computeSpindleForcesKernel (:385-420) and integrateNodesKernel
    (:525-571): a driven node displaced by d from spindle + offset feels
    f = k (target - x) + c (v_s - v); one step gives v = f dt / m, then
    x += v dt, then v *= exp(-0.001 dt) (damping 0.1 * 0.01).  A force
    beyond m 1e8 is scaled back to it (:537-545)."""
    k, c, dt = 5e7, 3.0, 1e-8
    T = mpmc.ToolOracle(spindle_stiffness=k, spindle_damping=c)
    T.set_mesh(single_tet())
    T.init_driven(2.0)                # threshold below the bottom: every node driven
    # |f| well below the integrator's cap m 1e8 (60 N for these nodes)
    ps, vs = np.array([1e-8, -2e-8, 3e-8]), np.array([0.1, 0.0, -0.2])
    T.set_spindle_state(ps, vs)
    before = T.tets()
    T.step(dt)
    after = T.tets()
    f = k * ps + c * vs               # target - x = ps (x at its original position), v = 0
    for i in range(4):
        v = f / before["mass"][i] * dt
        x = np.array([before["x"][i], before["y"][i], before["z"][i]]) + v * dt
        assert np.allclose([after["x"][i], after["y"][i], after["z"][i]], x, rtol=1e-15, atol=1e-18)
        assert np.allclose([after["vx"][i], after["vy"][i], after["vz"][i]], v * math.exp(-0.1 * 0.01 * dt),
                           rtol=1e-12)
    # the acceleration cap: a 3 um pull (150 N) on 6e-7 kg nodes is cut to 1e8 m/s^2
    T = mpmc.ToolOracle(spindle_stiffness=k, spindle_damping=0.0)
    T.set_mesh(single_tet())
    T.init_driven(2.0)
    T.set_spindle_state((3e-6, 0.0, 0.0), (0.0, 0.0, 0.0))
    T.step(dt)
    assert np.allclose(T.tets()["vx"], 1e8 * dt * math.exp(-0.1 * 0.01 * dt), rtol=1e-12)


def test_thermal_conduction_conserves_heat_and_uniform_is_steady():
    """thermalConductionTetKernel (:38-84) + applyTetHeatKernel (:87-97):
    sum_i Q_i = -k V dt sum_j (sum_i gradN_i) . gradN_j T_j = 0 (the shape
    gradients sum to zero), so sum m cp T is conserved; a uniform T makes
    no flux; the conduction runs on steps with step % 10 == 0 only."""
    m = toolfem.cylinder_tool_mesh(1.5e-3, 3e-3, H, base=(0, 0, 0))
    m["tet_region"][:] = toolfem.REGION_TOOL_BODY
    T = mpmc.ToolOracle(spindle_enabled=0, coolant_htc=0.0)
    T.set_mesh(m)
    s = T.surface()
    s["temperature"] = 25.0
    T.set_surface(s)
    T.step(1e-6)
    assert np.allclose(T.tets()["temperature"], 25.0, rtol=0, atol=1e-12)
    # heat one surface node: its excess goes to its tet node (fraction 0.35);
    # 45 degC keeps every node below the radiation threshold (35 degC)
    s = T.surface()
    s["temperature"][0] = 45.0
    T.set_surface(s)
    tets0 = T.tets()
    e0 = (tets0["mass"] * tets0["temperature"]).sum() * 200.0 + s["mass"][0] * 200.0 * s["temperature"][0]
    for _ in range(10):     # steps 1..10: the conduction runs at step 10
        T.step(1e-6)
    tets1, s1 = T.tets(), T.surface()
    e1 = (tets1["mass"] * tets1["temperature"]).sum() * 200.0 + s1["mass"][0] * 200.0 * s1["temperature"][0]
    # radiation: no node exceeds 25 + 10 degC; coolant off: heat conserved
    assert tets1["temperature"].max() < 35.0
    assert e1 == pytest.approx(e0, rel=1e-12)
    # the first transfer moved 0.35 of the excess (tet node at 25): 45 -> 38
    T2 = mpmc.ToolOracle(spindle_enabled=0, coolant_htc=0.0)
    T2.set_mesh(m)
    s = T2.surface()
    s["temperature"][0] = 45.0
    T2.set_surface(s)
    T2.step(1e-6)
    assert T2.surface()["temperature"][0] == pytest.approx(45.0 - 0.35 * 20.0, rel=1e-12)


def test_thermal_bc_shank_coolant_radiation():
    """applyThermalBCKernel (:112-165): SHANK nodes are clamped to 25; a
    body node at T cools by min(h A (T - 25) dt / (m cp), 0.5 (T - 25)) with
    A = (max(m / rho, 1e-15))^(2/3), then radiates 0.7 sigma A (Tk^4 - Ta^4)
    dt / (m cp) (capped at 0.3 (T - 25)) when still above 35 degC.  (No
    conduction: k = 0; no surface transfer: the surface stays at 25.)"""
    m = single_tet(1e-3)
    m["tet_region"] = np.array([toolfem.REGION_SHANK, 1, 1, 1], np.int32)
    htc, dt = 5e4, 1e-3
    T = mpmc.ToolOracle(spindle_enabled=0, coolant_htc=htc, thermal_conductivity=0.0)
    T.set_mesh(m)
    t = T.tets()
    t["temperature"] = [900.0, 400.0, 30.0, 36.0]
    T.set_tets(t)
    T.step(dt)
    got = T.tets()["temperature"]
    assert got[0] == 25.0
    mass = t["mass"]
    for i in (1, 2, 3):
        Ti = t["temperature"][i]
        area = max(mass[i] / 14500.0, 1e-15) ** (2.0 / 3.0)
        if Ti > 25.1:
            Ti -= min(htc * area * (Ti - 25.0) * dt / (mass[i] * 200.0), 0.5 * (Ti - 25.0))
        if Ti > 35.0:
            rad = 0.7 * 5.670374419e-8 * area * ((Ti + 273.15) ** 4 - 298.15 ** 4) * dt / (mass[i] * 200.0)
            Ti -= min(rad, 0.3 * (Ti - 25.0))
        assert got[i] == pytest.approx(Ti, rel=1e-12)


def test_wear_rate_closed_form():
    """updateWearKernel (:573-588): a contacting surface node gains
    1e-9 * vm * |v| * exp(-1000 / max(T + 273.15, 300)) * dt; beyond 0.3 mm
    it turns WORN."""
    T = mpmc.ToolOracle(spindle_enabled=0)
    T.set_mesh(single_tet())
    s = T.surface()
    s["in_contact"][1] = 1
    s["von_mises"][1] = 2e9
    s["vx"][1] = 3.0
    s["temperature"][1] = 500.0
    T.set_surface(s)
    T.step(1e-3)
    w = T.surface()
    # the surface T drops by the transfer first (0.35 of the excess over its tet node)
    Ts = 500.0 - 0.35 * 475.0
    want = 1e-9 * 2e9 * 3.0 * math.exp(-1000.0 / (Ts + 273.15)) * 1e-3
    assert w["wear"][1] == pytest.approx(want, rel=1e-12)
    assert w["wear"][0] == 0.0
    s = T.surface()
    s["wear"][1] = 0.31e-3
    T.set_surface(s)
    T.step(1e-9)
    assert T.surface()["status"][1] == mpmc.WORN


def test_hash_slots_and_cell_bounds():
    """computeNodeHashKernel (SharedSpatialHash.cu:9-31): slot = ((cx 73856093)
    ^ (cy 19349663) ^ (cz 83492791)) mod T with c = max(0, floor((x - o) /
    cell)); T = nextPrime(max(2 N, 100003)); the cell bounds
    (findToolCellBoundsKernel, ContactSolver.cu:123-137) cover every node
    once, in slot order, ties in node order."""
    assert mpmc.next_prime(100003) == 100003 and mpmc.next_prime(100004) == 100019
    o = np.array([-0.01, -0.01, -0.01])
    p = np.array([0.0012, -0.0031, 0.00049])
    cell = 2e-4
    c = np.floor((p - o) / cell).astype(np.int64)
    want = ((c[0] * 73856093) ^ (c[1] * 19349663) ^ (c[2] * 83492791)) % 100003
    assert mpmc.node_hash(p, o, cell, 100003) == want
    assert mpmc.node_hash(o - 1.0, o, cell, 100003) == 0      # negative cells clamp to 0
    # a tool and one particle far away: build the hash through a resolve
    m = toolfem.cylinder_tool_mesh(1e-3, 2e-3, H, base=(0, 0, 0))
    T = mpmc.ToolOracle(spindle_enabled=0)
    T.set_mesh(m)
    S = mpms.Oracle(mpms.default_params(domain_min=tuple(o), domain_max=(0.01, 0.01, 0.01)))
    a = mpms.new_particles(1)
    a["x"], a["y"], a["z"] = 0.009, 0.009, 0.009
    a["mass"], a["volume"], a["id"] = 1e-9, 1e-12, 0
    S.set_particles(a)
    Cn = mpmc.ContactOracle(contact_radius=cell)
    Cn.resolve(S, T, 1e-8)
    nn = T.counts()[0]
    start, end, order = Cn.hash_table(nn)
    assert len(start) == 100003
    sx = T.surface()
    slots = np.array([mpmc.node_hash((sx["x"][i], sx["y"][i], sx["z"][i]), o, cell, 100003) for i in range(nn)])
    assert np.array_equal(order, np.lexsort((np.arange(nn), slots)))
    covered = np.zeros(nn, int)
    for h in np.nonzero(start >= 0)[0]:
        seg = order[start[h]:end[h]]
        assert np.all(slots[seg] == h)
        covered[seg] += 1
    assert np.all(covered == 1)
    assert Cn.results()["contacts"] == 0


def _pair(d, v=(0.0, 0.0, 0.0), r=2e-4, tp=25.0, tn=25.0, a=0.05, **cp):
    """One particle at distance d along +x from the only surface node of a
    one-tet tool (edge a) within the contact radius."""
    m = single_tet(a)
    T = mpmc.ToolOracle(spindle_enabled=0)
    T.set_mesh(m)
    s = T.surface()
    s["temperature"][0] = tn
    T.set_surface(s)
    S = mpms.Oracle(mpms.default_params(domain_min=(-0.01, -0.01, -0.01), domain_max=(0.01, 0.01, 0.01)))
    a = mpms.new_particles(1)
    a["x"], a["y"], a["z"] = d, 0.0, 0.0
    a["vx"], a["vy"], a["vz"] = v
    a["mass"], a["volume"], a["id"] = 4.43e-9, 1e-12, 0
    a["temperature"] = tp
    S.set_particles(a)
    C = mpmc.ContactOracle(contact_radius=r, **cp)
    return T, S, C


def test_contact_normal_projection_and_hertz_closed_form():
    """xpbdContactSpatialHashKernel (:159-408) + applyNodeHeatAndHertz
    (:665-791) for one particle at d < r from one node, first call
    (engagement 0.1): x moves by (r - d) along the normal, the inward
    velocity is removed, f_n = 0.1 (k pen + 2 zeta sqrt(k m) |v_n|), the
    node's reaction is -f_n n, its von Mises 0.1 * 3 f_n / (2 pi r pen)."""
    d, r, vin = 1.2e-4, 2e-4, -0.7
    T, S, C = _pair(d, v=(vin, 0.0, 0.0), r=r)
    C.resolve(S, T, 1e-8)
    p = S.get_particles()[0]
    pen = r - d
    assert p["x"] == pytest.approx(d + pen, rel=1e-14)
    assert p["vx"] == pytest.approx(0.0, abs=1e-15)
    k, zeta, m = 1e7, 0.25, 4.43e-9
    fn = 0.1 * (k * pen + 2 * zeta * math.sqrt(k * m) * abs(vin))
    res = C.results()
    assert res["contacts"] == 1 and res["engagement"] == pytest.approx(0.1)
    assert res["force"] == pytest.approx(fn, rel=1e-12)
    nd = T.surface()[0]
    assert nd["fx"] == pytest.approx(-fn, rel=1e-12) and nd["in_contact"] == 1
    assert nd["von_mises"] == pytest.approx(0.1 * 3 * fn / (2 * math.pi * r * pen), rel=1e-12)
    assert nd["fn"] == 0.0 and nd["pen"] == 0.0      # reset after the Hertz pass
    assert p["temperature"] == 25.0                   # no sliding, equal temperatures: no heat


def test_contact_friction_heat_split_and_inertial_cap():
    """Sliding contact: the frictional heat f_t v_t dt splits into the node's
    share hp and the particle's 1 - hp, the conduction term moving heat from
    the hotter side and cancelling in the sum (node heat + particle heat =
    the generated heat); the friction reaction on the node points along the
    particle's relative tangential velocity; f_t <= m v_t / dt."""
    d, r, vt, dt = 1.0e-4, 2e-4, 4.0, 1e-8
    # tool node below the radiation threshold (ambient + 1), no conduction term
    T, S, C = _pair(d, v=(0.0, vt, 0.0), r=r, tp=300.0, tn=25.5, a=1e-3, contact_htc=0.0)
    C.resolve(S, T, dt)
    res = C.results()
    p = S.get_particles()[0]
    nd = T.surface()[0]
    # particle heating (its share, positive): dT = (1 - hp) Q / (rho_w V cp_w)
    Q = res["heat"]
    assert Q > 0
    dTp = (1 - 0.5) * Q / (7850.0 * 1e-12 * 475.0)
    assert p["temperature"] == pytest.approx(300.0 + dTp, rel=1e-9)
    # the node's share raised its temperature: hp Q / (m cp_tool)
    m_node = T.surface()["mass"][0]
    assert nd["temperature"] - 25.5 == pytest.approx(0.5 * Q / (m_node * 200.0), rel=1e-8)
    ft_y = nd["fy"]
    assert ft_y > 0                                   # dragged along the particle's motion
    assert ft_y <= 4.43e-9 * vt / dt * (1 + 1e-12)
    # generated heat f_t v_t dt times the engagement (0.1; the friction force
    # itself is not scaled)
    assert Q == pytest.approx(ft_y * vt * dt * 0.1, rel=1e-12)


def test_contact_skips_far_failed_inactive_and_fixed():
    """No contact at d >= r, with a FAILED node, or for INACTIVE and
    FIXED_BOUNDARY particles (:188-190, :221)."""
    for d, status, fail in ((2.0e-4, mpms.ACTIVE, False), (1e-4, mpms.INACTIVE, False),
                            (1e-4, mpms.FIXED, False), (1e-4, mpms.ACTIVE, True)):
        T, S, C = _pair(d)
        a = S.get_particles()
        a["status"] = status
        S.set_particles(a)
        if fail:
            s = T.surface()
            s["status"][0] = mpmc.FAILED
            T.set_surface(s)
        C.resolve(S, T, 1e-8)
        assert C.results()["contacts"] == 0
        assert S.get_particles()[0]["x"] == d


def test_engagement_ramp_and_explosion_guard():
    """Engagement (:1097-1112): after k calls with contact, scale = 0.1 +
    0.9 s^2 (3 - 2 s), s = k / 200.  The guard (:1276-1296): more contacts
    than 50 per node zero the results and the node accumulators."""
    T, S, C = _pair(1.5e-4)
    for k in range(5):
        a = S.get_particles()
        a["x"] = 1.5e-4
        S.set_particles(a)
        C.resolve(S, T, 1e-8)
        s = k / 200.0
        assert C.results()["engagement"] == pytest.approx(0.1 + 0.9 * s * s * (3 - 2 * s), rel=1e-14)
    # more contacts than 50 per surface node (the one-tet tool has 4): 201
    # particles on a sphere of radius 0.15 mm around node 0 (the others are
    # 5 cm away)
    def ring(n):
        a = mpms.new_particles(n)
        u = (np.arange(n) + 0.5) / n
        th, ph = np.arccos(1 - 2 * u), np.pi * (1 + 5 ** 0.5) * np.arange(n)
        a["x"], a["y"], a["z"] = 1.5e-4 * np.sin(th) * np.cos(ph), 1.5e-4 * np.sin(th) * np.sin(ph), 1.5e-4 * np.cos(th)
        a["mass"], a["volume"], a["id"] = 4.43e-9, 1e-12, np.arange(n)
        return a
    T, S, C = _pair(1.5e-4)
    S.set_particles(ring(201))
    C.resolve(S, T, 1e-8)
    assert C.results()["contacts"] == 0 and C.results()["force"] == 0.0
    assert T.surface()["in_contact"][0] == 0
    # 200 contacts are allowed
    T, S, C = _pair(1.5e-4)
    S.set_particles(ring(200))
    C.resolve(S, T, 1e-8)
    assert C.results()["contacts"] == 200
