"""Workpiece + tool set-ups of the benchmark configurations.

Material: Ti-6Al-4V-like Johnson-Cook set (A=1098 MPa, B=1092 MPa, n=0.93,
C=0.014, m=1.1, eps0_dot=1/s, Tr=293 K, Tm=1878 K), rho=4430 kg/m^3,
E=113.8 GPa, nu=0.342, cp=526 J/kgK.  Cell h = 0.25 mm, 8 particles per
cell with seeded jitter.
"""
import math

import numpy as np

H = 2.5e-4
SEED_SLAB = 0xED6E0001
SEED_DRILL = 0xED6E0002
SEED_BATCH = 0xED6E0003
SEED_SERVE = 0xED6E0004

MATERIAL = dict(
    density=4430.0, young=113.8e9, poisson=0.342, specific_heat=526.0,
    melt_temp=1878.0, ref_temp=293.0,
    jc_A=1098e6, jc_B=1092e6, jc_n=0.93, jc_C=0.014, jc_m=1.1, ref_strain_rate=1.0,
    failure_strain=0.3, damage_threshold=1.0, critical_plastic_strain=0.15,
    thermal_chip_ratio=0.25, contact_damping=0.15, max_contact_temp_rise=50.0,
    thermal_conductivity=42.6, kinematic_hardening=0.0, thermal_clamp_enabled=1,
    has_config=1, cell_size=H, tool_contact_radius=1.5 * H,
    min_time_step=1e-12, max_time_step=1e-4, max_particles=2 ** 31 - 1,
)


def params_dict(fixture=0.0, **kw):
    d = dict(MATERIAL)
    d["fixture_thickness"] = fixture
    d.update(kw)
    return d


def rake_face_samples(y0, y1, z_edge, x_edge=0.0, rake_deg=5.0, height=1.5e-3, spacing=H / 2):
    """Point samples of a flat rake-face patch: the plane through the cutting
    edge (x_edge, y, z_edge), leaning back by the rake angle, from the edge
    up to `height` along the face; cutting direction +x."""
    g = math.radians(rake_deg)
    up = np.array([-math.sin(g), 0.0, math.cos(g)])
    ny = int(round((y1 - y0) / spacing)) + 1
    ns = int(round(height / spacing)) + 1
    pts = []
    for j in range(ns):
        for i in range(ny):
            pts.append(np.array([x_edge, y0 + i * spacing, z_edge]) + j * spacing * up)
    return np.array(pts)


def slab_config0():
    """Config 0: 32x32x16 cells (131,072 particles), rake-face point tool at
    5 deg rake, 0.5 mm depth of cut; the workpiece moves at -1 m/s in the tool
    frame, fixture 0; dt = 1e-8 s, 200 steps."""
    cx, cy, cz = 32, 32, 16
    top = cz * H
    return dict(
        name="config0_rake_slab_32x32x16",
        cells=(cx, cy, cz), seed=SEED_SLAB, origin=(0.0, 0.0, 0.0),
        vel=(-1.0, 0.0, 0.0), temperature=293.0,
        params=params_dict(fixture=0.0),
        tool=("points", rake_face_samples(0.0, cy * H, top - 0.5e-3), 1.5 * H),
        tip=(0.0, 0.0, top - 0.5e-3),
        dt=1e-8, steps=200,
    )


def drill_desc_values():
    return dict(kind=3, band=0.0, friction=0.3, grad_eps=0.05 * H, rake_angle=0.0, clearance_angle=0.0,
                radius=5e-3, helix_angle=math.radians(30.0), point_angle=math.radians(118.0),
                flute_offset=3e-3, flute_radius=2.25e-3, flute_length=20e-3)


def drill_config1():
    """Config 1: 256x128x64 cells (16,777,216 particles), analytic two-flute
    helical drill (D=10 mm, 30 deg helix, 118 deg point) at 3000 rpm with
    0.1 mm/rev feed, tip 0.5 mm below the top face; dt = 1e-8 s, 1000 steps."""
    cx, cy, cz = 256, 128, 64
    top = cz * H
    omega = 3000.0 * 2.0 * math.pi / 60.0
    feed = 0.1e-3 * 3000.0 / 60.0
    return dict(
        name="config1_drill_slab_256x128x64",
        cells=(cx, cy, cz), seed=SEED_DRILL, origin=(0.0, 0.0, 0.0),
        vel=(0.0, 0.0, 0.0), temperature=293.0,
        params=params_dict(fixture=2e-3),
        tool=("drill", drill_desc_values(),
              dict(center=(cx * H / 2, cy * H / 2, top - 0.5e-3), angle=0.0, vel=(0.0, 0.0, -feed),
                   omega=omega)),
        tip=(cx * H / 2, cy * H / 2, top - 0.5e-3),
        dt=1e-8, steps=1000,
    )


def scaled(cfg, cells, steps=None, name=None):
    """The same set-up on a smaller slab (tool placed relative to the top face)."""
    c = dict(cfg)
    cx, cy, cz = cells
    top = cz * H
    c["cells"] = cells
    if steps is not None:
        c["steps"] = steps
    if cfg["tool"][0] == "points":
        c["tool"] = ("points", rake_face_samples(0.0, cy * H, top - 0.5e-3), cfg["tool"][2])
        c["tip"] = (0.0, 0.0, top - 0.5e-3)
    elif cfg["tool"][0] == "drill":
        pose = dict(cfg["tool"][2])
        pose["center"] = (cx * H / 2, cy * H / 2, top - 0.5e-3)
        c["tool"] = ("drill", cfg["tool"][1], pose)
        c["tip"] = pose["center"]
    c["name"] = name or f"{cfg['name']}@{cx}x{cy}x{cz}"
    return c


def wedge_desc_values(rake_deg=5.0, clearance_deg=7.0, friction=0.3):
    return dict(kind=2, band=0.0, friction=friction, grad_eps=0.05 * H, rake_angle=math.radians(rake_deg),
                clearance_angle=math.radians(clearance_deg), radius=0.0, helix_angle=0.0, point_angle=0.0,
                flute_offset=0.0, flute_radius=0.0, flute_length=0.0)


def prism_tool_mesh(width, length=1.5e-3, rake_deg=5.0, clearance_deg=7.0):
    """Triangle mesh of a straight cutting insert in its rest frame: the
    cutting edge on the y axis (|y| <= width/2), the rake face leaning back by
    the rake angle from the edge up to `length`, the flank face rising by the
    clearance angle behind the edge, a back face closing the section (the
    solid the analytic wedge describes, cut to a finite prism).  Returns
    (verts, tris, normals) with outward unit normals."""
    g, a = math.radians(rake_deg), math.radians(clearance_deg)
    edge = np.array([0.0, 0.0])
    rake = length * np.array([-math.sin(g), math.cos(g)])
    flank = length * np.array([-math.cos(a), math.sin(a)])
    sec = [edge, rake, flank]
    y0, y1 = -0.5 * width, 0.5 * width
    V = np.array([[p[0], y, p[1]] for y in (y0, y1) for p in sec])  # 0-2 at y0, 3-5 at y1
    quads = [(0, 3, 4, 1), (0, 2, 5, 3), (1, 4, 5, 2)]
    T = []
    for i, j, k, l in quads:
        T += [(i, j, k), (i, k, l)]
    T += [(0, 1, 2), (3, 5, 4)]
    T = np.array(T, dtype=np.int32)
    centroid = V.mean(axis=0)
    N = []
    for t in T:
        n = np.cross(V[t[1]] - V[t[0]], V[t[2]] - V[t[0]])
        n /= np.linalg.norm(n)
        if np.dot(n, V[t].mean(axis=0) - centroid) < 0:
            n = -n
        N.append(n)
    return V, T, np.array(N)


def voxel_desc_values(friction=0.3):
    return dict(kind=4, band=0.0, friction=friction, grad_eps=0.05 * H, rake_angle=0.0, clearance_angle=0.0,
                radius=0.0, helix_angle=0.0, point_angle=0.0, flute_offset=0.0, flute_radius=0.0,
                flute_length=0.0)


def voxel_tool_case(cfg):
    """A wedge workpiece of the batched distribution cut by the insert mesh
    instead of the analytic wedge: voxel SDF built with the reference's
    padding of 2 mm and 128 voxels along the longest axis (ContactSolver.cu:1019),
    the insert spanning the workpiece width plus a cell each side."""
    c = dict(cfg)
    cx, cy, cz = cfg["cells"]
    pose = dict(cfg["tool"][2])
    ex, _, ez = pose["center"]
    pose["center"] = (ex, 0.5 * cy * H, ez)
    mesh = prism_tool_mesh(cy * H + 2 * H)
    c["tool"] = ("voxel", voxel_desc_values(), pose, mesh, 0.002, 128)
    c["tip"] = pose["center"]
    c["name"] = cfg["name"].replace("batch", "voxel")
    return c


def sample_workpiece(rng, index, seed, steps):
    """One workpiece of the batched / served distributions: N log-uniform in
    [2e4, 2e6] particles (cells = N/8 in a 2:2:1 box), J-C A, B, n, C, m each
    uniform within +-15 % of config 0, an analytic 5 deg rake wedge entering at
    the -x face with cutting speed uniform in [0.5, 5] m/s and depth of cut
    uniform in [0.1, 1.0] mm, fixture 0.5 mm, dt = 1e-8 s.  The cutting edge
    starts one cell inside the -x face, so the rake face is in contact from
    the first step (a 500-step run moves the tool by at most 25 um)."""
    n = math.exp(rng.uniform(math.log(2e4), math.log(2e6)))
    a = (n / 32.0) ** (1.0 / 3.0)
    cx, cy, cz = max(2, int(round(2 * a))), max(2, int(round(2 * a))), max(2, int(round(a)))
    jc = {k: MATERIAL[k] * rng.uniform(0.85, 1.15) for k in ("jc_A", "jc_B", "jc_n", "jc_C", "jc_m")}
    speed = rng.uniform(0.5, 5.0)
    doc = min(rng.uniform(0.1e-3, 1.0e-3), 0.6 * cz * H)
    top = cz * H
    edge = (H, 0.0, top - doc)
    return dict(
        name=f"batch_{index:03d}_{cx}x{cy}x{cz}",
        cells=(cx, cy, cz), seed=(seed * 0x9E3779B97F4A7C15 + index) & 0xFFFFFFFFFFFFFFFF,
        origin=(0.0, 0.0, 0.0), vel=(0.0, 0.0, 0.0), temperature=293.0,
        params=params_dict(fixture=0.5e-3, **jc),
        tool=("wedge", wedge_desc_values(), dict(center=edge, angle=0.0, vel=(speed, 0.0, 0.0), omega=0.0)),
        tip=edge, dt=1e-8, steps=steps, speed=speed, depth_of_cut=doc,
    )


def batch_config2(count=64, seed=SEED_BATCH, steps=500):
    """Config 2: `count` independent workpieces of the batched distribution
    (sample_workpiece), 500 steps each."""
    rng = np.random.default_rng(seed)
    return [sample_workpiece(rng, i, seed, steps) for i in range(count)]


def apply_to_solver(mpmw, solver, cfg):
    """Load a configuration into an mpmw.Solver."""
    cx, cy, cz = cfg["cells"]
    solver.init_jittered_slab(cfg["seed"], cx, cy, cz, cfg["origin"], cfg["vel"], cfg["temperature"])
    kind = cfg["tool"][0]
    if kind == "points":
        solver.set_tool_samples(cfg["tool"][1], cfg["tool"][2])
    elif kind in ("drill", "wedge"):
        d = mpmw.ToolDesc()
        for k, v in cfg["tool"][1].items():
            setattr(d, k, v)
        solver.set_tool_analytic(d)
        p = mpmw.ToolPose()
        pd = cfg["tool"][2]
        p.center[:] = pd["center"]
        p.angle = pd["angle"]
        p.vel[:] = pd["vel"]
        p.omega = pd["omega"]
        solver.set_tool_pose(p)
    elif kind == "voxel":
        d = mpmw.ToolDesc()
        for k, v in cfg["tool"][1].items():
            setattr(d, k, v)
        verts, tris, normals = cfg["tool"][3]
        solver.set_tool_sdf_mesh(d, verts, tris, normals, cfg["tool"][4], cfg["tool"][5])
        p = mpmw.ToolPose()
        pd = cfg["tool"][2]
        p.center[:] = pd["center"]
        p.angle = pd["angle"]
        p.vel[:] = pd["vel"]
        p.omega = pd["omega"]
        solver.set_tool_pose(p)
    solver.set_tool_tip(*cfg["tip"])
    if cfg.get("transfer") == "apic":
        solver.set_transfer(mpmw.TRANSFER_APIC)
    elif cfg.get("transfer") == "apic_textbook":
        solver.set_transfer(mpmw.TRANSFER_APIC_TEXTBOOK)
