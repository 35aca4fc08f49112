"""ctypes binding of the CPU FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product library never does.  The oracle restates
src/MPMWorkpieceSolver.cu of the reference in C (see mpm_oracle.c).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liborc_mpm.so")

ACTIVE, INACTIVE, CHIP, BOUNDARY, FIXED_BOUNDARY = 0, 1, 2, 3, 4
TOOL_NONE, TOOL_POINTS, TOOL_WEDGE, TOOL_DRILL, TOOL_VOXEL = 0, 1, 2, 3, 4
TRANSFER_PIC, TRANSFER_APIC, TRANSFER_APIC_TEXTBOOK = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "density", "young", "poisson", "specific_heat", "melt_temp", "ref_temp",
        "jc_A", "jc_B", "jc_n", "jc_C", "jc_m", "ref_strain_rate",
        "failure_strain", "damage_threshold", "critical_plastic_strain", "thermal_chip_ratio",
        "tool_contact_radius", "contact_damping", "max_contact_temp_rise",
        "thermal_conductivity", "kinematic_hardening")] + [
        ("thermal_clamp_enabled", C.c_int32), ("has_config", C.c_int32),
        ("cell_size", C.c_double), ("particle_spacing_factor", C.c_double),
        ("fixture_thickness", C.c_double), ("min_time_step", C.c_double),
        ("max_time_step", C.c_double), ("max_particles", C.c_int32), ("pad_", C.c_int32)]


class ToolDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32)] + [(n, C.c_double) for n in (
        "band", "friction", "grad_eps", "rake_angle", "clearance_angle", "radius",
        "helix_angle", "point_angle", "flute_offset", "flute_radius", "flute_length")]


class ToolPose(C.Structure):
    _fields_ = [("center", C.c_double * 3), ("angle", C.c_double),
                ("vel", C.c_double * 3), ("omega", C.c_double)]


class Grid(C.Structure):
    _fields_ = [("cell", C.c_double), ("minx", C.c_double), ("miny", C.c_double),
                ("minz", C.c_double), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32)]


class Arrays(C.Structure):
    _fields_ = [("n", C.c_int32), ("pad_", C.c_int32)] + [(n, _dp) for n in (
        "pos", "vel", "fext", "mass", "volume", "density", "temperature",
        "plastic_strain", "strain_rate", "damage", "residual_stress", "stress",
        "backstress")] + [("id", _ip), ("status", _ip)]


class Metrics(C.Structure):
    _fields_ = [("max_stress", C.c_double), ("max_temperature", C.c_double),
                ("kinetic_energy", C.c_double), ("contact_force", C.c_double),
                ("time", C.c_double), ("step", C.c_int64)]


class Particle(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz", "mass", "volume", "density",
        "temperature", "plasticStrain", "strainRate", "damage", "residualStress")] + [
        ("s", C.c_double * 6), ("a", C.c_double * 6), ("id", C.c_int32), ("status", C.c_int32),
        ("C", C.c_double * 9)]


_lib = None


class Sdf(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("pad_", C.c_int32),
                ("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("voxel", C.c_float),
                ("grid", C.POINTER(C.c_float))]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Params)]
        L.orc_destroy.argtypes = [C.c_void_p]
        for name in ("orc_set_particles", "orc_get_particles"):
            getattr(L, name).argtypes = [C.c_void_p, C.POINTER(Arrays)]
            getattr(L, name).restype = C.c_int
        L.orc_particle_count.argtypes = [C.c_void_p]
        L.orc_particle_count.restype = C.c_int
        L.orc_step.argtypes = [C.c_void_p, C.c_double]
        L.orc_run.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_stable_dt.argtypes = [C.c_void_p]
        L.orc_stable_dt.restype = C.c_double
        L.orc_get_metrics.argtypes = [C.c_void_p, C.POINTER(Metrics)]
        L.orc_sync_metrics.argtypes = [C.c_void_p]
        L.orc_compact.argtypes = [C.c_void_p]
        L.orc_round_state_fp32.argtypes = [C.c_void_p]
        L.orc_replace_particles.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_set_fp32_emulation.argtypes = [C.c_void_p, C.c_int]
        L.orc_reset.argtypes = [C.c_void_p]
        L.orc_get_bounds.argtypes = [C.c_void_p, _dp]
        L.orc_temperature_at.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.orc_temperature_at.restype = C.c_double
        L.orc_apply_heat_flux.argtypes = [C.c_void_p] + [C.c_double] * 4
        L.orc_set_particle_temperatures.argtypes = [C.c_void_p, _dp, C.c_int]
        L.orc_set_tool_samples.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double]
        L.orc_set_tool_analytic.argtypes = [C.c_void_p, C.POINTER(ToolDesc)]
        L.orc_set_tool_pose.argtypes = [C.c_void_p, C.POINTER(ToolPose)]
        L.orc_set_tool_tip.argtypes = [C.c_void_p] + [C.c_double] * 3
        L.orc_apply_external_force.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.orc_init_particle_box.argtypes = [C.c_void_p, _dp, _dp, C.c_double]
        L.orc_init_cylinder.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_get_grid.argtypes = [C.c_void_p, C.POINTER(Grid)]
        L.orc_grid_node_count.argtypes = [C.c_void_p]
        L.orc_grid_node_count.restype = C.c_int
        L.orc_get_grid_snapshot.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_get_grid_snapshot.restype = C.c_int
        L.orc_get_grid_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _ip]
        L.orc_get_grid_state.restype = C.c_int
        L.orc_weights.argtypes = [C.POINTER(Grid), C.c_double, C.c_double, C.c_double, _ip, _dp, _dp]
        L.orc_equivalent_stress.argtypes = [C.c_double] * 6
        L.orc_equivalent_stress.restype = C.c_double
        L.orc_jc_flow.argtypes = [C.POINTER(Params), C.c_double, C.c_double, C.c_double]
        L.orc_jc_flow.restype = C.c_double
        L.orc_rebuild_params.argtypes = [C.c_double, _dp, _dp, C.POINTER(Grid)]
        L.orc_chip_kick.argtypes = [C.c_int32, _dp]
        L.orc_tool_phi.argtypes = [C.POINTER(ToolDesc), C.POINTER(ToolPose)] + [C.c_double] * 3
        L.orc_tool_phi.restype = C.c_double
        L.orc_tool_normal.argtypes = [C.POINTER(ToolDesc), C.POINTER(ToolPose)] + [C.c_double] * 3 + [_dp]
        L.orc_constitutive.argtypes = [C.POINTER(Params), C.POINTER(Particle), _dp, C.c_double]
        L.orc_params_default.argtypes = [C.POINTER(Params)]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_gen_jittered_slab.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_double, _dp,
                                            C.c_double, C.c_double, _dp, _dp, _dp, _ip, _ip]
        _fp = C.POINTER(C.c_float)
        L.orc_point_tri_d2.argtypes = [_fp, _fp, _fp, _fp, _fp]
        L.orc_point_tri_d2.restype = C.c_float
        L.orc_sdf_build.argtypes = [_dp, C.c_int, _ip, _dp, C.c_int, C.c_double, C.c_int, C.POINTER(Sdf)]
        L.orc_sdf_free.argtypes = [C.POINTER(Sdf)]
        L.orc_sdf_sample.argtypes = [C.POINTER(Sdf), _dp]
        L.orc_sdf_sample.restype = C.c_double
        L.orc_sdf_in_range.argtypes = [C.POINTER(Sdf), _dp]
        L.orc_set_tool_sdf.argtypes = [C.c_void_p, C.POINTER(Sdf), C.POINTER(ToolDesc)]
        L.orc_set_transfer.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_affine.argtypes = [C.c_void_p, _dp]
        L.orc_get_affine.restype = C.c_int
        L.orc_set_affine.argtypes = [C.c_void_p, _dp]
        L.orc_weights_quadratic.argtypes = [C.POINTER(Grid), C.c_double, C.c_double, C.c_double, _ip, _dp, _dp,
                                            _dp]
        _lib = L
    return _lib


def _p(a, ct=_dp):
    return None if a is None else a.ctypes.data_as(ct)


def default_params(**kw):
    p = Params()
    lib().orc_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def make_arrays(n, with_fext=True, with_backstress=True):
    """Allocate a dict of numpy arrays for n particles (FP64, AoS-free)."""
    d = {
        "pos": np.zeros((n, 3)), "vel": np.zeros((n, 3)),
        "fext": np.zeros((n, 3)) if with_fext else None,
        "mass": np.zeros(n), "volume": np.zeros(n), "density": np.zeros(n),
        "temperature": np.zeros(n), "plastic_strain": np.zeros(n), "strain_rate": np.zeros(n),
        "damage": np.zeros(n), "residual_stress": np.zeros(n),
        "stress": np.zeros((n, 6)), "backstress": np.zeros((n, 6)) if with_backstress else None,
        "id": np.zeros(n, dtype=np.int32), "status": np.zeros(n, dtype=np.int32),
    }
    return d


def _to_struct(d):
    a = Arrays()
    a.n = len(d["mass"])
    for k in ("pos", "vel", "fext", "mass", "volume", "density", "temperature",
              "plastic_strain", "strain_rate", "damage", "residual_stress", "stress", "backstress"):
        v = d.get(k)
        if v is not None:
            assert v.dtype == np.float64 and v.flags.c_contiguous, k
        setattr(a, k, _p(v))
    for k in ("id", "status"):
        assert d[k].dtype == np.int32 and d[k].flags.c_contiguous, k
        setattr(a, k, _p(d[k], _ip))
    return a


def gen_jittered_slab(seed, cx, cy, cz, h, origin=(0.0, 0.0, 0.0), density=4430.0, fixture=0.0):
    n = cx * cy * cz * 8
    d = make_arrays(n)
    d["density"][:] = density
    org = np.asarray(origin, dtype=np.float64)
    lib().orc_gen_jittered_slab(seed, cx, cy, cz, h, _p(org), density, fixture,
                                _p(d["pos"]), _p(d["mass"]), _p(d["volume"]),
                                _p(d["id"], _ip), _p(d["status"], _ip))
    return d


class Oracle:
    def __init__(self, params=None, **kw):
        self.params = params if params is not None else default_params(**kw)
        self.h = lib().orc_create(C.byref(self.params))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_particles(self, d):
        self._keep = d
        return lib().orc_set_particles(self.h, C.byref(_to_struct(d)))

    def replace_particles(self, d):
        """Replace the particle state; grid, step count and time are kept."""
        a = _to_struct(d)
        lib().orc_replace_particles(self.h, C.byref(a))

    def count(self):
        return lib().orc_particle_count(self.h)

    def get_particles(self):
        d = make_arrays(self.count())
        a = _to_struct(d)
        lib().orc_get_particles(self.h, C.byref(a))
        return d

    def init_box(self, mn, mx, spacing):
        lib().orc_init_particle_box(self.h, _p(np.asarray(mn, float)), _p(np.asarray(mx, float)), spacing)

    def init_cylinder(self, c, radius, length, spacing, axis=2):
        lib().orc_init_cylinder(self.h, _p(np.asarray(c, float)), radius, length, spacing, axis)

    def set_tool_samples(self, xyz, contact_radius):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        lib().orc_set_tool_samples(self.h, _p(xyz), len(xyz), contact_radius)

    def set_tool_analytic(self, desc):
        lib().orc_set_tool_analytic(self.h, C.byref(desc))

    def set_tool_sdf(self, sdf, desc):
        """Voxel tool: `sdf` from build_sdf (its grid is copied)."""
        lib().orc_set_tool_sdf(self.h, C.byref(sdf.raw), C.byref(desc))

    def set_tool_pose(self, pose):
        lib().orc_set_tool_pose(self.h, C.byref(pose))

    def set_transfer(self, mode):
        """TRANSFER_PIC (trilinear), TRANSFER_APIC (quadratic, affine C = 0) or
        TRANSFER_APIC_TEXTBOOK (quadratic with C = 4/h^2 B and no decay)."""
        lib().orc_set_transfer(self.h, mode)

    def get_affine(self):
        C9 = np.zeros((self.count(), 9))
        lib().orc_get_affine(self.h, _p(C9))
        return C9

    def set_affine(self, C9):
        C9 = np.ascontiguousarray(C9, dtype=np.float64).reshape(self.count(), 9)
        lib().orc_set_affine(self.h, _p(C9))

    def set_tool_tip(self, x, y, z):
        lib().orc_set_tool_tip(self.h, x, y, z)

    def apply_external_force(self, c, radius, f):
        lib().orc_apply_external_force(self.h, _p(np.asarray(c, float)), radius, _p(np.asarray(f, float)))

    def step(self, dt):
        lib().orc_step(self.h, dt)

    def run(self, nsteps, dt):
        lib().orc_run(self.h, nsteps, dt)

    def stable_dt(self):
        return lib().orc_stable_dt(self.h)

    def metrics(self):
        m = Metrics()
        lib().orc_get_metrics(self.h, C.byref(m))
        return m

    def sync_metrics(self):
        lib().orc_sync_metrics(self.h)

    def compact(self):
        lib().orc_compact(self.h)

    def set_fp32_emulation(self, on):
        """Sensitivity probe: FP32-rounded P2G sums and stored state."""
        lib().orc_set_fp32_emulation(self.h, 1 if on else 0)

    def round_state_fp32(self):
        """Round every stored particle field but the positions to FP32."""
        lib().orc_round_state_fp32(self.h)

    def reset(self):
        lib().orc_reset(self.h)

    def bounds(self):
        b = np.zeros(6)
        lib().orc_get_bounds(self.h, _p(b))
        return b

    def temperature_at(self, x, y, z):
        return lib().orc_temperature_at(self.h, x, y, z)

    def apply_heat_flux(self, x, y, z, q):
        lib().orc_apply_heat_flux(self.h, x, y, z, q)

    def set_particle_temperatures(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        lib().orc_set_particle_temperatures(self.h, _p(t), len(t))

    def grid(self):
        g = Grid()
        lib().orc_get_grid(self.h, C.byref(g))
        return g

    def grid_snapshot(self):
        n = lib().orc_grid_node_count(self.h)
        m, p, f, tm = np.zeros(n), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
        lib().orc_get_grid_snapshot(self.h, _p(m), _p(p), _p(f), _p(tm))
        return {"mass": m, "mom": p, "force": f, "tm": tm}

    def grid_state(self):
        n = lib().orc_grid_node_count(self.h)
        m, v, tm = np.zeros(n), np.zeros((n, 3)), np.zeros(n)
        fx = np.zeros(n, dtype=np.int32)
        lib().orc_get_grid_state(self.h, _p(m), _p(v), _p(tm), _p(fx, _ip))
        return {"mass": m, "vel": v, "tm": tm, "fixed": fx}


def weights(grid, x, y, z):
    idx = np.zeros(8, dtype=np.int32)
    w = np.zeros(8)
    g = np.zeros(24)
    lib().orc_weights(C.byref(grid), x, y, z, _p(idx, _ip), _p(w), _p(g))
    return idx, w, g.reshape(8, 3)


def weights_quadratic(grid, x, y, z):
    """The 27-node quadratic stencil: (idx, w, grad[27, 3], dpos[27, 3])."""
    idx = np.zeros(27, dtype=np.int32)
    w = np.zeros(27)
    g = np.zeros(81)
    d = np.zeros(81)
    lib().orc_weights_quadratic(C.byref(grid), x, y, z, _p(idx, _ip), _p(w), _p(g), _p(d))
    return idx, w, g.reshape(27, 3), d.reshape(27, 3)


def equivalent_stress(*s):
    return lib().orc_equivalent_stress(*[float(v) for v in s])


def jc_flow(params, eps_p, rate, temp):
    return lib().orc_jc_flow(C.byref(params), eps_p, rate, temp)


def rebuild_params(cell, mn, mx):
    g = Grid()
    lib().orc_rebuild_params(cell, _p(np.asarray(mn, float)), _p(np.asarray(mx, float)), C.byref(g))
    return g


def chip_kick(pid):
    r = np.zeros(3)
    lib().orc_chip_kick(pid, _p(r))
    return r


def tool_phi(desc, pose, x, y, z):
    return lib().orc_tool_phi(C.byref(desc), C.byref(pose), x, y, z)


def tool_normal(desc, pose, x, y, z):
    n = np.zeros(3)
    lib().orc_tool_normal(C.byref(desc), C.byref(pose), x, y, z, _p(n))
    return n


def constitutive(params, particle, L, dt):
    L = np.ascontiguousarray(L, dtype=np.float64).reshape(9)
    lib().orc_constitutive(C.byref(params), C.byref(particle), _p(L), dt)
    return particle


def set_threads(n):
    lib().orc_set_threads(n)


def point_tri_d2(p, a, b, c):
    """The reference's squared point-triangle distance and closest point (float)."""
    f = lambda v: np.ascontiguousarray(v, dtype=np.float32)
    p, a, b, c = f(p), f(a), f(b), f(c)
    q = np.zeros(3, dtype=np.float32)
    fp = C.POINTER(C.c_float)
    d2 = lib().orc_point_tri_d2(p.ctypes.data_as(fp), a.ctypes.data_as(fp), b.ctypes.data_as(fp),
                                c.ctypes.data_as(fp), q.ctypes.data_as(fp))
    return float(d2), q


def vertex_normals(nverts, tris, normals):
    """ContactSolver::initSDF (src/ContactSolver.cu:1024-1044): each face
    normal divided by its length when that exceeds 1e-12, added to its three
    vertices; each vertex sum divided by its length the same way.  FP64, in
    triangle order."""
    vn = np.zeros((nverts, 3))
    for t, tri in enumerate(np.asarray(tris)):
        n = [float(c) for c in normals[t]]
        ln = np.sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])
        if ln > 1e-12:
            n = [c / ln for c in n]
        for j in tri:
            for k in range(3):
                vn[j, k] += n[k]
    for i in range(nverts):
        v = vn[i]
        ln = np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
        if ln > 1e-12:
            vn[i] = [c / ln for c in v]
    return vn


def worn_vertices(verts, tris, normals, wear):
    """ContactSolver::morphToolSDF (src/ContactSolver.cu:1053-1075): vertex i
    moved to rest - normal * wear[i] (no wear past the array's end)."""
    v = np.asarray(verts, dtype=np.float64)
    vn = vertex_normals(len(v), tris, normals)
    w = np.zeros(len(v))
    k = min(len(wear), len(v))
    w[:k] = np.asarray(wear, dtype=np.float64)[:k]
    return v - vn * w[:, None]


class SdfGrid:
    """A voxel SDF built by the oracle (owns its grid)."""

    def __init__(self, verts, tris, normals, padding=0.002, voxels_per_axis=128):
        v = np.ascontiguousarray(verts, dtype=np.float64)
        t = np.ascontiguousarray(tris, dtype=np.int32)
        n = np.ascontiguousarray(normals, dtype=np.float64)
        self.raw = Sdf()
        rc = lib().orc_sdf_build(_p(v), len(v), _p(t, _ip), _p(n), len(t), padding, voxels_per_axis,
                                 C.byref(self.raw))
        if rc != 0:
            raise ValueError("orc_sdf_build: bad mesh")

    @property
    def dims(self):
        return (self.raw.nx, self.raw.ny, self.raw.nz)

    @property
    def origin(self):
        return (self.raw.ox, self.raw.oy, self.raw.oz)

    @property
    def voxel(self):
        return self.raw.voxel

    def grid(self):
        """Values as an (nz, ny, nx) float32 array."""
        nx, ny, nz = self.dims
        return np.ctypeslib.as_array(self.raw.grid, shape=(nz * ny * nx,)).reshape(nz, ny, nx).copy()

    def sample(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        return lib().orc_sdf_sample(C.byref(self.raw), _p(q))

    def in_range(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        return bool(lib().orc_sdf_in_range(C.byref(self.raw), _p(q)))

    def __del__(self):
        try:
            lib().orc_sdf_free(C.byref(self.raw))
        except Exception:
            pass
