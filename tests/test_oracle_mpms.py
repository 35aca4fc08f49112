"""Pins of the F-tracking oracle (oracle/mpms_oracle.c, the restatement of
MPMSolver::step, src/MPMSolver.cu) by closed forms, invariants and
hand-derived cases, on the CPU.  Each test names the reference lines it
pins; the known answers are derived here from the mathematics, not retyped
from the oracle."""
import math

import numpy as np
import pytest

import mpms

MU = lambda p: p.young / (2 * (1 + p.poisson))                                  # noqa: E731
LAM = lambda p: p.young * p.poisson / ((1 + p.poisson) * (1 - 2 * p.poisson))    # noqa: E731


def one(x=(0.0, 0.0, 0.0), **kw):
    a = mpms.new_particles(1)
    a["x"], a["y"], a["z"] = x
    a["mass"] = 1e-7
    a["volume"] = 2e-11
    a["density"] = 4430.0
    a["id"] = 0
    for k, v in kw.items():
        a[k] = v
    return a


def test_quadratic_weights_partition_gradient_and_linear_reproduction():
    """getWeightsAndGradients (:19-34): sum w = 1, sum dw = 0, sum w x_i = x_p
    and sum dw x_i = 1 at random positions; a particle's 27 node masses sum to
    its mass and its momentum to m v (grid of one particle, no forces)."""
    rng = np.random.default_rng(1)
    p = mpms.default_params(lod_enabled=0, fixture_thickness=0.0)
    for _ in range(20):
        x = rng.uniform(-0.01, 0.01, 3)
        # a heavy particle: every node it reaches holds more than the G2P's
        # 1e-12 kg emptiness threshold (:538)
        a = one(tuple(x), vx=0.3, vy=-0.2, vz=0.1, mass=1.0)
        a["status"] = mpms.BOUNDARY          # no stress force
        o = mpms.Oracle(p)
        o.set_particles(a)
        o.step(1e-12)
        g = o.grid()
        assert g["mass"].sum() == pytest.approx(1.0, rel=1e-13)
        # momentum after the update: m v + gravity m g dt
        assert g["mom"].sum(axis=0) == pytest.approx(np.array([0.3, -0.2, 0.1 - 9.81e-12]), rel=1e-12)
        # the gathered velocity equals the particle's (partition of unity)
        b = o.get_particles()[0]
        assert (b["vx"], b["vy"]) == pytest.approx((0.3, -0.2), rel=1e-12)


def test_polar_decomposition_rotation_and_stretch():
    """Higham's iteration: a rotation is its own polar factor (R = F), and
    F = Q S with S symmetric positive definite gives R = Q, also for F with
    det F = 1e-3 and condition number 1e6."""
    rng = np.random.default_rng(3)
    for _ in range(10):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        if np.linalg.det(q) < 0:
            q[:, 0] = -q[:, 0]
        R, it = mpms.polar(q)
        assert np.abs(R - q).max() < 1e-14 and it <= 2
        m = rng.normal(size=(3, 3))
        S = m @ m.T + 0.5 * np.eye(3)
        R, _ = mpms.polar(q @ S)
        assert np.abs(R - q).max() < 1e-12
        u, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        S = u @ np.diag([10.0, 10.0, 1e-5]) @ u.T
        R, it = mpms.polar(q @ S)
        assert np.abs(R - q).max() < 1e-9 and np.abs(R.T @ R - np.eye(3)).max() < 1e-12 and it < 30


def _stress(p, F=np.eye(3), Fp=np.eye(3), **kw):
    a = one(**kw)
    a["F"] = np.asarray(F, float).reshape(9)
    a["Fp"] = np.asarray(Fp, float).reshape(9)
    return mpms.p2g_stress(p, a[0])


def test_neo_hookean_uniaxial_and_clamp():
    """Compressible Neo-Hookean of p2gKernel (:171-196), elastic (A huge):
    F = diag(l, 1, 1): sigma_xx = mu/J (l^2 - 1) + lambda (J - 1), sigma_yy =
    lambda (J - 1) with J = l; J is clamped to [0.9, 1.1] (:174-175) and
    Fe = F Fp^-1."""
    p = mpms.default_params(jc_A=1e15)
    mu, lam = MU(p), LAM(p)
    for l in (1.01, 0.95, 1.05):
        s, J, skip, _ = _stress(p, F=np.diag([l, 1, 1]))
        assert J == pytest.approx(l) and not skip
        assert s[0, 0] == pytest.approx(mu / l * (l * l - 1) + lam * (l - 1), rel=1e-12)
        assert s[1, 1] == pytest.approx(lam * (l - 1), rel=1e-12)
        assert abs(s[0, 1]) < 1e-3
    s, J, _, _ = _stress(p, F=np.diag([1.3, 1, 1]))
    assert J == 1.1 and s[1, 1] == pytest.approx(lam * 0.1, rel=1e-12)
    # Fe = F Fp^-1: the same F with Fp = F is stress-free
    F = np.array([[1.02, 0.01, 0], [0, 0.99, 0.02], [0.01, 0, 1.0]])
    s, J, _, _ = _stress(p, F=F, Fp=F)
    assert np.abs(s).max() < 1e-2 * 1e-9 * mu + 1e-3


def test_corotated_uniaxial_and_rotation_invariance():
    """Fixed-corotated extension: F = diag(l, 1, 1) has R = I and sigma_xx =
    2 mu / J (l - 1) l + lambda (J - 1); a rigid rotation Q F rotates the
    stress to Q sigma Q^T."""
    p = mpms.default_params(jc_A=1e15, elastic_model=mpms.ELASTIC_COROTATED)
    mu, lam = MU(p), LAM(p)
    l = 1.03
    s, J, _, _ = _stress(p, F=np.diag([l, 1, 1]))
    assert s[0, 0] == pytest.approx(2 * mu / l * (l - 1) * l + lam * (l - 1), rel=1e-12)
    assert s[1, 1] == pytest.approx(lam * (l - 1), rel=1e-12)
    th = 0.7
    Q = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    F = np.array([[1.02, 0.01, 0.0], [0.0, 0.98, 0.015], [0.005, 0.0, 1.01]])
    s0, _, _, _ = _stress(p, F=F)
    s1, _, _, _ = _stress(p, F=Q @ F)
    assert np.abs(s1 - Q @ s0 @ Q.T).max() < 1e-6 * np.abs(s0).max()


def test_radial_return_closed_form():
    """J-C radial return (:214-276): after the return the shifted deviator's
    von Mises value is the flow stress Y (evaluated at the old eps_p), the
    increment is (q - Y) / 3 mu, the pressure is kept, heat 0.9 Y d /
    (rho cp), det Fp = 1, and the increment is capped at 0.05."""
    p = mpms.default_params(jc_A=300e6, jc_B=0.0, jc_C=0.0, lod_enabled=0)
    mu = MU(p)
    F = np.diag([1.004, 0.998, 0.998])
    s_tr, _, _, _ = _stress(mpms.default_params(jc_A=1e15), F=F)
    dev = s_tr - np.trace(s_tr) / 3 * np.eye(3)
    q = math.sqrt(1.5 * (dev * dev).sum())
    assert q > 300e6
    s, J, _, a = _stress(p, F=F, temperature=25.0)
    devn = s - np.trace(s) / 3 * np.eye(3)
    assert math.sqrt(1.5 * (devn * devn).sum()) == pytest.approx(300e6, rel=1e-12)
    assert np.trace(s) == pytest.approx(np.trace(s_tr), rel=1e-12)
    d = (q - 300e6) / (3 * mu)
    assert a["plastic_strain"] == pytest.approx(d, rel=1e-12)
    assert a["temperature"] == pytest.approx(25.0 + 0.9 * 300e6 * d / (p.density * p.specific_heat), rel=1e-12)
    assert np.linalg.det(a["Fp"].reshape(3, 3)) == pytest.approx(1.0, abs=1e-14)
    # cap: a strongly over-stressed state takes 0.05
    s, J, _, a = _stress(p, F=np.diag([1.2, 0.95, 0.95]))
    assert a["plastic_strain"] == pytest.approx(0.05, rel=1e-15)


def test_newton_return_linear_hardening_closed_form():
    """The consistent (Newton) return extension: with Y = A + B eps (n = 1,
    C = 0, T* = 0) the increment is (q - A - B eps0) / (3 mu + B) and the
    returned von Mises value is A + B (eps0 + d); with n < 1 the residual
    q - 3 mu d - Y(eps0 + d) vanishes."""
    A, B, e0 = 300e6, 2e9, 0.01
    p = mpms.default_params(jc_A=A, jc_B=B, jc_n=1.0, jc_C=0.0, return_map=mpms.RETURN_NEWTON, lod_enabled=0)
    mu = MU(p)
    F = np.diag([1.004, 0.998, 0.998])
    s_tr, _, _, _ = _stress(mpms.default_params(jc_A=1e15), F=F)
    dev = s_tr - np.trace(s_tr) / 3 * np.eye(3)
    q = math.sqrt(1.5 * (dev * dev).sum())
    s, _, _, a = _stress(p, F=F, plastic_strain=e0)
    d = (q - A - B * e0) / (3 * mu + B)
    assert a["plastic_strain"] - e0 == pytest.approx(d, rel=1e-12)
    devn = s - np.trace(s) / 3 * np.eye(3)
    assert math.sqrt(1.5 * (devn * devn).sum()) == pytest.approx(A + B * (e0 + d), rel=1e-12)
    p2 = mpms.default_params(jc_A=A, jc_B=B, jc_n=0.3, jc_C=0.02, return_map=mpms.RETURN_NEWTON, lod_enabled=0)
    F2 = np.diag([1.012, 0.994, 0.994])
    s_tr, _, _, _ = _stress(mpms.default_params(jc_A=1e15), F=F2)
    dev = s_tr - np.trace(s_tr) / 3 * np.eye(3)
    q = math.sqrt(1.5 * (dev * dev).sum())
    assert q > (A + B * e0 ** 0.3) * (1 + 0.02 * math.log(1e4))      # plastic
    s, _, _, a = _stress(p2, F=F2, plastic_strain=e0, strain_rate=1e4)
    d = a["plastic_strain"] - e0
    assert 0 < d < 0.05
    Y = (A + B * (e0 + d) ** 0.3) * (1 + 0.02 * math.log(1e4))
    assert abs(q - 3 * mu * d - Y) <= 1e-9 * q


def test_damage_cdm_and_chip_transition():
    """JC failure (:313-340): D += d / eps_f with eps_f = (D1 + D2 exp(D3
    sigma*)) (1 + D4 ln max(rate*, 1)) (1 + D5 T*), sigma* = p / q (trial);
    past 0.7 of the threshold the stress scales by (Dth - D) / (0.3 Dth);
    at the threshold the particle becomes CHIP and adds no stress force."""
    p = mpms.default_params(jc_A=300e6, jc_B=0.0, jc_C=0.0, lod_enabled=0)
    F = np.diag([1.004, 0.998, 0.998])
    s_tr, _, _, _ = _stress(mpms.default_params(jc_A=1e15), F=F)
    dev = s_tr - np.trace(s_tr) / 3 * np.eye(3)
    q = math.sqrt(1.5 * (dev * dev).sum())
    tri = np.trace(s_tr) / 3 / q
    ef = (0.06 + 3.31 * math.exp(-1.96 * tri)) * 1.0 * 1.0
    d = (q - 300e6) / (3 * MU(p))
    s, _, skip, a = _stress(p, F=F, damage=0.5)
    assert a["damage"] == pytest.approx(0.5 + d / ef, rel=1e-12) and not skip
    D = 0.8
    s, _, skip, a = _stress(p, F=F, damage=D - d / ef)
    g = (1.0 - a["damage"]) / 0.3
    devn = s - np.trace(s) / 3 * np.eye(3)
    assert math.sqrt(1.5 * (devn * devn).sum()) == pytest.approx(g * 300e6, rel=1e-9)
    s, _, skip, a = _stress(p, F=F, damage=0.9999999)
    assert a["status"] == mpms.CHIP and skip


def test_gravity_free_fall_and_kill_zone():
    """updateGridKernel (:424-442): a lone particle away from the fixture
    falls with g = -9.81 (v_z = g k dt after k steps, x_z = sum); a particle
    in the outer cell layer is made INACTIVE by the P2G (:131-139)."""
    p = mpms.default_params(lod_enabled=0, fixture_thickness=0.0)
    o = mpms.Oracle(p)
    a = one((0.001, 0.002, 0.003))
    a["status"] = mpms.BOUNDARY
    o.set_particles(a)
    dt = 1e-6
    for _ in range(10):
        o.step(dt)
    b = o.get_particles()[0]
    assert b["vz"] == pytest.approx(-9.81 * 10 * dt, rel=1e-10)
    assert b["z"] == pytest.approx(0.003 - 9.81 * dt * dt * 55, rel=1e-9)
    edge = one((-0.1 + 0.3 * p.dx, 0.0, 0.0))
    o2 = mpms.Oracle(p)
    o2.set_particles(edge)
    o2.step(1e-8)
    assert o2.get_particles()[0]["status"] == mpms.INACTIVE


def test_fixture_damping_and_clamp():
    """applyFixtureBCKernel (:81-111): within the fixture layer above the
    domain's bottom, node velocity is damped by 1 / (1 + c (1 - zr)^2 dt);
    the bottom node layer is clamped to zero."""
    p = mpms.default_params(lod_enabled=0, fixture_thickness=0.004, domain_min=(-0.01, -0.01, 0.0),
                            domain_max=(0.01, 0.01, 0.02))
    dt = 1e-8
    # a particle whose stencil spans node layers 2..4 (z = 1.5 dx .. ): all in the fixture zone
    z = 3.0 * p.dx
    a = one((0.0, 0.0, z), vx=1.0)
    a["status"] = mpms.BOUNDARY
    o = mpms.Oracle(p)
    o.set_particles(a)
    o.step(dt)
    # gathered vx = sum_k w_k vx_k with each layer k damped by its own factor
    b = o.get_particles()[0]
    f = lambda k: 1.0 / (1.0 + 1e7 * (1.0 - k * p.dx / 0.004) ** 2 * dt)     # noqa: E731
    xf = z / p.dx
    base = math.floor(xf - 0.5)
    d0 = xf - base
    wz = [0.5 * (1.5 - d0) ** 2, 0.75 - (d0 - 1) ** 2, 0.5 * (d0 - 0.5) ** 2]
    want = sum(w * f(base + k) for k, w in enumerate(wz))
    assert b["vx"] == pytest.approx(want, rel=1e-12)


def test_tool_rasterisation_is_sticky():
    """rasterizeToolKernel + applyGridCollisionKernel (:444-492): nodes the
    tool marks take the tool node's velocity, so a particle whose 27 nodes
    are all marked moves with the tool."""
    p = mpms.default_params(lod_enabled=0, fixture_thickness=0.0)
    a = one((0.0011, 0.0013, 0.0017), mass=1.0)   # every node above the 1e-12 kg threshold
    a["status"] = mpms.BOUNDARY
    o = mpms.Oracle(p)
    o.set_particles(a)
    g = np.arange(-2, 4) * p.dx
    pts = np.array([(0.001 + x, 0.001 + y, 0.0015 + z) for x in g for y in g for z in g])
    o.set_tool_nodes(pts, np.tile([0.5, -1.0, 2.0], (len(pts), 1)))
    o.step(1e-8)
    b = o.get_particles()[0]
    assert (b["vx"], b["vy"], b["vz"]) == pytest.approx((0.5, -1.0, 2.0), rel=1e-12)


def test_lod_zones_far_particles_stress_free():
    """lodClassifyKernel (:655-676): ACTIVE within the active radius, NEAR
    within the near radius, FAR beyond; a FAR particle adds no stress force
    (:143-146) and its C is zeroed in G2P (:580-586)."""
    p = mpms.default_params(fixture_thickness=0.0)
    a = mpms.new_particles(3)
    for i, r in enumerate((0.001, 0.005, 0.02)):
        a[i]["x"] = r
        a[i]["mass"], a[i]["volume"], a[i]["id"] = 1e-7, 2e-11, i
        a[i]["F"] = np.diag([1.01, 1.0, 1.0]).reshape(9)
        a[i]["C"] = np.full(9, 3.0)
    o = mpms.Oracle(p)
    o.set_tool_position(0.0, 0.0, 0.0)
    o.set_particles(a)
    o.step(1e-8)
    b = o.get_particles()
    assert list(b["lod"]) == [mpms.LOD_ACTIVE, mpms.LOD_NEAR, mpms.LOD_FAR]
    assert np.all(b[2]["C"] == 0.0) and np.all(b[2]["s"] == 0.0)
    assert np.abs(b[0]["s"]).max() > 0


def test_apic_affine_field_and_F_update():
    """An affine velocity v = v0 + A x on a particle block with C = A: the
    gathered velocity of an interior particle is v0 + A x_p, its C becomes
    0.99 * 3/dx^2 * (dx^2 / 4) A = 0.7425 A (:577-591), F becomes
    (I + A dt) F (:594-601) and the strain rate sqrt(2/3 (sum A_ii^2 + 1/2
    sum_(i<j) (A_ij + A_ji)^2))."""
    p = mpms.default_params(lod_enabled=0, fixture_thickness=0.0, jc_A=1e15)
    dx = p.dx
    A = np.array([[1.0, 2.0, -0.5], [0.3, -1.5, 0.7], [-0.2, 0.4, 0.5]]) * 1e3
    v0 = np.array([0.2, -0.1, 0.3])
    xs = np.arange(-4, 5) * 0.5 * dx + 0.013 * dx
    pts = np.array([(x, y, z) for x in xs for y in xs for z in xs])
    a = mpms.new_particles(len(pts))
    a["x"], a["y"], a["z"] = pts.T
    a["mass"], a["volume"], a["id"] = 1e-7, 2e-11, np.arange(len(pts))
    a["status"] = mpms.BOUNDARY
    v = v0 + pts @ A.T
    a["vx"], a["vy"], a["vz"] = v.T
    a["C"] = A.reshape(9)
    o = mpms.Oracle(p)
    o.set_particles(a)
    dt = 1e-9
    o.step(dt)
    b = o.get_particles()
    i = int(np.argmin(np.abs(pts).sum(axis=1)))
    want_v = v0 + pts[i] @ A.T + np.array([0, 0, -9.81 * dt])
    assert np.array([b[i]["vx"], b[i]["vy"], b[i]["vz"]]) == pytest.approx(want_v, rel=1e-9, abs=1e-12)
    assert b[i]["C"].reshape(3, 3) == pytest.approx(0.7425 * A, rel=1e-8)
    assert b[i]["F"].reshape(3, 3) == pytest.approx(np.eye(3) + A * dt, rel=1e-9)
    S = A + A.T
    rate = math.sqrt(2 / 3 * (np.trace(A * A * np.eye(3)) + 0.5 * (S[0, 1] ** 2 + S[0, 2] ** 2 + S[1, 2] ** 2)))
    assert b[i]["strain_rate"] == pytest.approx(rate, rel=1e-9)


def test_chip_pass_and_external_force():
    """chipFractureAndSeparationKernel of MPMSolver (:1007-1060, even
    steps): ACTIVE below tip - 1.5 dx with damage > 1e-6 becomes CHIP with
    its stress zeroed and residual max(res, |diag sigma|); a CHIP with damage
    >= 1.5 Dth is kicked by the u32 hash of its id and reset to 0.5 Dth.
    applyExternalForceKernel (:607-627): f (1 - d / r) replaces the force of
    ACTIVE and CHIP particles within r."""
    p = mpms.default_params(lod_enabled=0, fixture_thickness=0.0, jc_A=1e15)
    a = mpms.new_particles(3)
    a["mass"], a["volume"] = 1e-7, 2e-11
    a["id"] = [0, 77, 2]
    a[0]["z"], a[0]["damage"] = -0.002, 1e-5
    a[0]["s"] = [3e6, -4e6, 12e6, 1e6, 0, 0]
    a[1]["status"], a[1]["damage"], a[1]["x"] = mpms.CHIP, 1.6, 0.003
    a[2]["x"], a[2]["status"] = 0.006, mpms.BOUNDARY
    o = mpms.Oracle(p)
    o.set_tool_position(0.0, 0.0, 0.0)
    o.set_particles(a)
    o.step(1e-12)
    b = o.get_particles()
    assert b[0]["status"] == mpms.CHIP and np.all(b[0]["s"] == 0) and b[0]["damage"] == 1.0
    seed = (77 * 2654435761) & 0xFFFFFFFF
    r = np.array([(seed >> 16) & 0xFF, (seed >> 8) & 0xFF, seed & 0xFF]) / 255.0 - 0.5
    assert b[1]["damage"] == 0.5
    assert np.array([b[1]["vx"], b[1]["vy"], b[1]["vz"] + 9.81e-12]) == pytest.approx(0.5 * r, rel=1e-9, abs=1e-15)
    o2 = mpms.Oracle(p)
    c = mpms.new_particles(2)
    c["mass"], c["volume"], c["x"] = 1e-7, 2e-11, [0.001, 0.0025]
    c[1]["status"] = mpms.BOUNDARY
    o2.set_particles(c)
    o2.apply_external_force((0.0, 0.0, 0.0), 0.002, (1.0, 2.0, -3.0))
    d = o2.get_particles()
    assert np.array([d[0]["efx"], d[0]["efy"], d[0]["efz"]]) == pytest.approx([0.5, 1.0, -1.5])
    assert d[1]["efx"] == 0.0


def test_lattice_init_and_stable_dt_and_metrics():
    """initializeParticleBox (:1290-1325): z, y, x ascending, the bottom
    tenth FIXED_BOUNDARY, mass rho * scaling * s^3; getStableTimeStep
    (:1170-1178) = 0.2 dx / sqrt(E / (rho * scaling)); metrics
    (computeParticleStatsKernel :726-785) by hand on two particles."""
    p = mpms.default_params(mass_scaling=4.0)
    o = mpms.Oracle(p)
    o.init_box((0, 0, 0), (1e-3, 5e-4, 1e-3), 2.5e-4)
    a = o.get_particles()
    assert len(a) == 5 * 3 * 5
    assert a[0]["z"] == 0.0 and a[1]["x"] == 2.5e-4 and a[5]["y"] == 2.5e-4 and a[15]["z"] == 2.5e-4
    assert np.all(a["status"][a["z"] <= 1e-4] == mpms.FIXED) and np.all(a["status"][a["z"] > 1e-4] == mpms.ACTIVE)
    assert a[0]["mass"] == pytest.approx(4430 * 4 * 2.5e-4 ** 3, rel=1e-15)
    assert o.stable_dt() == pytest.approx(0.2 * p.dx / math.sqrt(p.young / (4430 * 4)), rel=1e-15)
    o.init_cylinder((0, 0, 0), 5e-4, 1e-3, 2.5e-4)
    c = o.get_particles()
    assert np.all(c["x"] ** 2 + c["y"] ** 2 <= 25e-8 + 1e-18) and c["z"].max() == 0.0
    m = mpms.new_particles(2)
    m["mass"] = [1.0, 2.0]
    m["vx"] = [1.0, 0.5]
    m["s"][0] = [200e6, 0, 0, 0, 0, 0]
    m["temperature"] = [300.0, 500.0]
    m["status"] = [mpms.ACTIVE, mpms.CHIP]
    o.set_particles(m)
    mt = o.metrics()
    assert mt["max_stress"] == pytest.approx(200e6) and mt["max_temperature"] == 500.0
    assert mt["kinetic_energy"] == pytest.approx(0.5 + 0.25)
