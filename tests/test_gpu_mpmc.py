"""GPU parity of the tool FEM step and the spatial-hash XPBD contact
(include/mpmc.h) against the FP64 oracle restatement (oracle/mpmc_oracle.c:
FEMSolver::step, src/FEMSolver.cu; ContactSolver's hash path,
src/ContactSolver.cu; SharedSpatialHash.cu).

Workloads: a tetrahedralised cylinder tool (toolfem.cylinder_tool_mesh)
driven down through spindle springs while it spins, with contact forces and
heat applied to surface nodes; a block of F-tracking particles the tool's
tip penetrates, sliding and hotter than the tool; the coupled step (tool
step, tool nodes to the MPM solver, MPM step, contact) over 30 steps.

Tolerances (normwise relative, parity.normwise): FP64 1e-9 on every field;
the contact with FP32 particles 1e-5 on positions, velocities, temperatures
and the node forces and heat.  The spindle damper is 1 N s/m: the
reference's default 1e4 N s/m is unstable under its explicit update when
c dt / m > 2 (a 1.5e-7 kg node at dt = 1e-8: 667).
"""
import numpy as np
import pytest

import ftrack
import mpmc
import mpms
import parity
import toolfem

pytestmark = pytest.mark.gpu

H = 5e-4


def tool_pair(mesh, **kw):
    """The oracle and the GPU tool on one mesh with the same parameters."""
    o = mpmc.ToolOracle(**kw)
    o.set_mesh(mesh)
    g = toolfem.Tool(toolfem.tool_params(**kw))
    g.set_mesh(mesh)
    return o, g


def node_errors(gd, od):
    """Per-field normwise errors of a GPU node dict against oracle nodes."""
    pos = np.stack([od["x"], od["y"], od["z"]], 1)
    vel = np.stack([od["vx"], od["vy"], od["vz"]], 1)
    frc = np.stack([od["fx"], od["fy"], od["fz"]], 1)
    e = {"pos": parity.normwise(gd["pos"], pos), "vel": parity.normwise(gd["vel"], vel),
         "force": parity.normwise(gd["force"], frc), "temperature": parity.normwise(gd["temperature"], od["temperature"]),
         "von_mises": parity.normwise(gd["von_mises"], od["von_mises"]), "wear": parity.normwise(gd["wear"], od["wear"]),
         "stress": parity.normwise(gd["stress"], od["s"])}
    assert np.array_equal(gd["status"], od["status"])
    return e


def _report(tag, errs):
    print(f"\n[{tag}] " + " ".join(f"{k}={v:.2e}" for k, v in sorted(errs.items())), flush=True)


@pytest.mark.parametrize("textbook", [0, 1])
def test_tool_fem_matches_oracle(textbook):
    """60 steps of the dynamic tool: spindle moving down at 5 m/s, the tool
    spinning at 314 rad/s about z, contact forces and heat on three surface
    nodes, the conduction on every tenth step."""
    mesh = toolfem.cylinder_tool_mesh(1.5e-3, 4e-3, H, base=(0.0, 0.0, 0.0))
    kw = dict(spindle_damping=1.0, polar_textbook=textbook)
    o, g = tool_pair(mesh, **kw)
    for q in (o, g):
        q.init_driven(0.2)
        q.set_transform((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 314.16)
        for i, f in ((3, (0.2, -0.1, 0.5)), (17, (-0.3, 0.0, 0.1)), (40, (0.0, 0.4, -0.2))):
            q.apply_contact_force(i, f, 2e-6)
    dt = 1e-8
    for k in range(60):
        for q in (o, g):
            q.set_spindle_state((0.0, 0.0, -5.0 * dt * k), (0.0, 0.0, -5.0))
            q.step(dt)
    et = node_errors(g.tets(), o.tets())
    es = node_errors(g.surface(), o.surface())
    _report(f"tool tets polar_textbook={textbook}", et)
    _report(f"tool surface polar_textbook={textbook}", es)
    ot = o.tets()
    assert np.abs(ot["z"] - ot["oz"]).max() > 1e-6            # the tool moved
    assert ot["temperature"].max() > 25.0 + 1e-9               # and conducted the applied heat
    for k, v in {**et, **{"s_" + a: b for a, b in es.items()}}.items():
        assert v <= 1e-9, (k, v)
    g.close()


def test_tool_thermal_subcycles_match_oracle():
    """Kinematic tool (no mechanics), dt = 5 ms: the conduction sub-cycles
    (alpha dt_thermal / h^2 > 0.25), the surface-to-tet transfer, the coolant
    and radiation boundary terms and the SHANK clamp, 25 steps."""
    mesh = toolfem.cylinder_tool_mesh(1.5e-3, 4e-3, H, base=(0.0, 0.0, 0.0))
    o, g = tool_pair(mesh, spindle_enabled=0, coolant_htc=2e4)
    ns = len(mesh["surf_xyz"])
    temp = 25.0 + 600.0 * (np.arange(ns) % 7 == 0)
    so = o.surface()
    so["temperature"] = temp
    o.set_surface(so)
    g.set_surface(temperature=temp)
    for _ in range(25):
        o.step(5e-3)
        g.step(5e-3)
    et, es = node_errors(g.tets(), o.tets()), node_errors(g.surface(), o.surface())
    _report("tool thermal tets", et)
    _report("tool thermal surface", es)
    assert o.tets()["temperature"].max() > 30.0
    for k in ("temperature",):
        assert et[k] <= 1e-9 and es[k] <= 1e-9, (et[k], es[k])


def particle_block(o):
    o.init_box((0, 0, 0), (3e-3, 3e-3, 2e-3), 2.5e-4)
    a = o.get_particles()
    a["id"] = np.arange(len(a))
    return a


def contact_setup(precision, steps_before=0):
    """Block 3 x 3 x 2 mm of particles at 0.25 mm (sliding in y at 0.3 m/s,
    at 300 degC, 10 % plastic strain) and the tool's tip 0.12 mm inside its
    top face."""
    mp = dict(lod_enabled=0, fixture_thickness=0.0, domain_min=(-0.01, -0.01, -0.01),
              domain_max=(0.01, 0.01, 0.01))
    so = mpms.Oracle(mpms.default_params(**mp))
    a = particle_block(so)
    a["vy"] = 0.3
    a["vz"] = -0.05
    a["temperature"] = 300.0
    a["plastic_strain"] = 0.1
    a["strain_rate"] = 2e3
    so.set_particles(a)
    sg = ftrack.Solver(ftrack.default_params(**mp, precision=ftrack.FP64 if precision == "fp64" else ftrack.FP32))
    import test_gpu_ftrack as tf
    sg.set_particles(tf.oracle_view(a))
    mesh = toolfem.cylinder_tool_mesh(1.0e-3, 3e-3, H, base=(1.5e-3, 1.5e-3, 2e-3 - 1.2e-4), face_samples=4)
    to, tg = tool_pair(mesh, spindle_damping=1.0)
    cp = dict(contact_radius=2e-4, contact_htc=1e5)
    co = mpmc.ContactOracle(**cp)
    cg = toolfem.Contact(toolfem.contact_params(**cp))
    return so, sg, to, tg, co, cg, tf


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_contact_matches_oracle(precision, tol):
    """Three resolves (the engagement ramp advancing) on the same particles
    and tool: particle positions, velocities and temperatures, the nodes'
    reaction forces, contact flags, Hertz von Mises and temperatures, and
    the totals."""
    so, sg, to, tg, co, cg, tf = contact_setup(precision)
    first = None
    for call in range(3):
        co.resolve(so, to, 1e-8)
        cg.resolve(sg, tg, 1e-8)
        ro, rg = co.results(), cg.results()
        print(f"\n[contact {precision} call {call}] oracle={ro} gpu={rg}", flush=True)
        first = ro["contacts"] if first is None else first
        assert rg["contacts"] == ro["contacts"]
        assert 0 < rg["probed"] < len(so.get_particles()) and rg["table_size"] == 100003
        for k in ("heat", "force", "cumulative_heat", "engagement"):
            assert abs(rg[k] - ro[k]) <= tol * abs(ro[k]), (k, rg[k], ro[k])
    assert first > 10            # the first call pushes the penetrated layer out
    gp, op = sg.get_particles(), tf.oracle_view(so.get_particles())
    errs, mism = tf.compare(gp, op)
    es = node_errors(tg.surface(), to.surface())
    assert np.array_equal(tg.surface()["in_contact"], to.surface()["in_contact"])
    _report(f"contact particles {precision}", errs)
    _report(f"contact nodes {precision}", es)
    assert mism == 0
    start = particle_block(mpms.Oracle(mpms.default_params(domain_min=(-0.01, -0.01, -0.01),
                                                            domain_max=(0.01, 0.01, 0.01))))
    ids = np.argsort(op["id"])
    moved = np.abs(op["pos"][ids] - np.stack([start[c] for c in "xyz"], 1)).max()
    assert moved > 1e-6                                      # the contact pushed particles
    gated = ("pos", "vel", "temperature") if precision == "fp32" else tuple(errs)
    for k in gated:
        assert errs[k] <= tol, (k, errs[k])
    for k in ("force", "temperature", "von_mises"):
        assert es[k] <= tol, (k, es[k])


def test_contact_explosion_guard_on_gpu():
    """More than 50 contacts per surface node zero the results and the node
    accumulators (:1276-1296); 200 on a 4-node tool are allowed."""
    import test_oracle_mpmc as to_
    for n, want in ((201, 0), (200, 200)):
        mesh = to_.single_tet(0.05)
        mp = dict(lod_enabled=0, fixture_thickness=0.0, domain_min=(-0.1, -0.1, -0.1), domain_max=(0.1, 0.1, 0.1))
        a = mpms.new_particles(n)
        u = (np.arange(n) + 0.5) / n
        th, ph = np.arccos(1 - 2 * u), np.pi * (1 + 5 ** 0.5) * np.arange(n)
        a["x"], a["y"], a["z"] = 1.5e-4 * np.sin(th) * np.cos(ph), 1.5e-4 * np.sin(th) * np.sin(ph), 1.5e-4 * np.cos(th)
        a["mass"], a["volume"], a["id"] = 4.43e-9, 1e-12, np.arange(n)
        import test_gpu_ftrack as tf
        sg = ftrack.Solver(ftrack.default_params(**mp, precision=ftrack.FP64))
        sg.set_particles(tf.oracle_view(a))
        tg = toolfem.Tool(toolfem.tool_params(spindle_enabled=0))
        tg.set_mesh(mesh)
        cg = toolfem.Contact(toolfem.contact_params(contact_radius=2e-4))
        cg.resolve(sg, tg, 1e-8)
        assert cg.results()["contacts"] == want
        assert tg.surface()["in_contact"][0] == (1 if want else 0)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_coupled_step_matches_oracle(precision, tol):
    """The engine's order (SimulationEngine.cpp:483-523, 611-638) for 30
    steps: tool step, the MPM solver's tool nodes := the tool's surface
    nodes, MPM step (sticky rasterised tool), contact resolve; the tool
    pressed down at 2 m/s through its spindle."""
    so, sg, to, tg, co, cg, tf = contact_setup(precision)
    for q in (to, tg):
        q.init_driven(0.2)
    dt = 1e-8
    most = 0
    for k in range(30):
        for q in (to, tg):
            q.set_spindle_state((0.0, 0.0, -2.0 * dt * k), (0.0, 0.0, -2.0))
            q.step(dt)
        s = to.surface()
        so.set_tool_nodes(np.stack([s["x"], s["y"], s["z"]], 1), np.stack([s["vx"], s["vy"], s["vz"]], 1))
        toolfem.link_tool_nodes(sg, tg)
        for q in (so, sg):
            q.set_tool_position(1.5e-3, 1.5e-3, 2e-3)
        so.step(dt)
        sg.step(dt)
        co.resolve(so, to, dt)
        cg.resolve(sg, tg, dt)
        assert cg.results()["contacts"] == co.results()["contacts"]
        most = max(most, co.results()["contacts"])
    gp, op = sg.get_particles(), tf.oracle_view(so.get_particles())
    errs, mism = tf.compare(gp, op)
    es = node_errors(tg.surface(), to.surface())
    et = node_errors(tg.tets(), to.tets())
    _report(f"coupled particles {precision}", errs)
    _report(f"coupled surface {precision}", es)
    _report(f"coupled tets {precision}", et)
    print(f"[coupled {precision}] oracle={co.results()} gpu={cg.results()}", flush=True)
    assert mism == 0 and most > 10
    gated = ("pos", "plastic_strain", "temperature") if precision == "fp32" else tuple(errs)
    for k in gated:
        assert errs[k] <= tol, (k, errs[k])
    for k in ("pos", "temperature"):
        assert es[k] <= tol and et[k] <= tol, (k, es[k], et[k])
