"""ctypes binding of libmpmw_hip.so (the C ABI in include/mpmw.h).

Python-side access to the MI355X MPM workpiece solver for tests and the
benchmark.  The product path is the HIP library; there is no CPU fallback:
loading fails loudly when the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmpmw_hip.so")

ACTIVE, INACTIVE, CHIP, BOUNDARY, FIXED_BOUNDARY = 0, 1, 2, 3, 4
OK, E_INVAL, E_HIP, E_STATE, E_NOMEM, E_RANGE = 0, -1, -2, -3, -4, -5
FP32, FP64 = 0, 1
TRANSFER_PIC, TRANSFER_APIC, TRANSFER_APIC_TEXTBOOK = 0, 1, 2
TOOL_NONE, TOOL_POINTS, TOOL_WEDGE, TOOL_DRILL, TOOL_VOXEL = 0, 1, 2, 3, 4
K_G2P_P2G, K_P2G, K_GRID, K_DIFFUSE, K_SORT, K_REDUCE, K_OTHER = range(7)
TIMING_OFF, TIMING_ALL, TIMING_PARTICLE_PASS = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

# Every symbol the C ABI exports (include/mpmw.h).
EXPORTS = (
    "mpmw_last_error", "mpmw_abi_version", "mpmw_params_default", "mpmw_create", "mpmw_destroy",
    "mpmw_reset", "mpmw_set_particles", "mpmw_get_particles", "mpmw_particle_count",
    "mpmw_init_particle_box", "mpmw_init_cylinder", "mpmw_gen_jittered_slab", "mpmw_init_jittered_slab",
    "mpmw_set_tool_samples", "mpmw_set_tool_analytic", "mpmw_set_tool_pose", "mpmw_set_tool_tip",
    "mpmw_apply_external_force", "mpmw_step", "mpmw_run", "mpmw_synchronize", "mpmw_stable_dt",
    "mpmw_get_metrics", "mpmw_sync_metrics", "mpmw_compact", "mpmw_get_bounds", "mpmw_temperature_at",
    "mpmw_apply_heat_flux", "mpmw_set_particle_temperatures", "mpmw_set_deterministic",
    "mpmw_get_grid_info", "mpmw_set_grid_snapshot", "mpmw_get_grid_snapshot", "mpmw_set_kernel_timing",
    "mpmw_get_kernel_times", "mpmw_get_stream", "mpmw_device_bytes", "mpmw_get_diagnostics",
    "mpmw_batch_create", "mpmw_batch_destroy", "mpmw_batch_add", "mpmw_batch_remove", "mpmw_batch_count",
    "mpmw_batch_step", "mpmw_batch_synchronize", "mpmw_set_tool_sdf_mesh", "mpmw_get_tool_sdf",
    "mpmw_set_transfer", "mpmw_get_affine", "mpmw_set_affine", "mpmw_export_ndjson", "mpmw_export_vtk",
    "mpmw_export_ndjson_noalloc", "mpmw_morph_tool_sdf", "mpmw_set_graph", "mpmw_graph_stats",
    "mpmw_batch_set_graph", "mpmw_batch_graph_stats", "mpmw_tool_mesh_params_default", "mpmw_parametric_tool_mesh",
    "mpmw_set_autotune", "mpmw_set_launch_config", "mpmw_get_tuning", "mpmw_batch_set_autotune",
    "mpmw_batch_set_launch_config", "mpmw_batch_get_tuning",
)


TOOLMESH_DRILL, TOOLMESH_BORING, TOOLMESH_REAMING, TOOLMESH_THREADING, TOOLMESH_MILLING = range(5)
TOOLMESH_NODES, TOOLMESH_TRIS = 674, 1344


class ToolMeshParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32), ("boring_bar_length_mm", C.c_double),
                ("boring_bar_diameter_mm", C.c_double), ("reamer_diameter_mm", C.c_double),
                ("thread_major_diameter_mm", C.c_double), ("thread_form_angle_deg", C.c_double)]


def parametric_tool_mesh(kind=TOOLMESH_DRILL, **over):
    """The reference application's parametric tool mesh (verts, tris,
    normals) of `kind`, with machining-parameter overrides (host only)."""
    p = ToolMeshParams()
    lib().mpmw_tool_mesh_params_default(C.byref(p))
    p.kind = kind
    for k, v in over.items():
        setattr(p, k, v)
    V = np.zeros((TOOLMESH_NODES, 3))
    T = np.zeros((TOOLMESH_TRIS, 3), dtype=np.int32)
    N = np.zeros((TOOLMESH_TRIS, 3))
    nv, nt = C.c_int32(), C.c_int32()
    _check(lib().mpmw_parametric_tool_mesh(C.byref(p), _p(V), _p(T, _ip), _p(N), C.byref(nv), C.byref(nt)),
           "parametric_tool_mesh")
    return V[:nv.value], T[:nt.value], N[:nt.value]


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
        ("max_time_step", C.c_double), ("max_particles", C.c_int32), ("precision", C.c_int32),
        ("deterministic", C.c_int32), ("device", C.c_int32)]


class ParticleArrays(C.Structure):
    _fields_ = [("n", C.c_int32), ("pad_", C.c_int32)] + [(n, _dp) for n in (
        "pos", "vel", "fext", "mass", "volume", "density", "temperature",
        "plastic_strain", "strain_rate", "damage", "residual_stress", "stress",
        "backstress")] + [("id", _ip), ("status", _ip)]


class ToolDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32)] + [(n, C.c_double) for n in (
        "band", "friction", "grad_eps", "rake_angle", "clearance_angle", "radius",
        "helix_angle", "point_angle", "flute_offset", "flute_radius", "flute_length")]


class ToolPose(C.Structure):
    _fields_ = [("center", C.c_double * 3), ("angle", C.c_double),
                ("vel", C.c_double * 3), ("omega", C.c_double)]


class Metrics(C.Structure):
    _fields_ = [("max_stress", C.c_double), ("max_temperature", C.c_double),
                ("kinetic_energy", C.c_double), ("contact_force", C.c_double),
                ("time", C.c_double), ("step", C.c_int64)]


class GridInfo(C.Structure):
    _fields_ = [("cell", C.c_double), ("minx", C.c_double), ("miny", C.c_double),
                ("minz", C.c_double), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("pad_", C.c_int32)]


class KernelTimes(C.Structure):
    _fields_ = [("ms", C.c_double * 8), ("launches", C.c_int64 * 8)]


class Diagnostics(C.Structure):
    _fields_ = [("sorts", C.c_int64), ("out_of_window", C.c_int64), ("steps_since_rebuild", C.c_int64),
                ("tiles", C.c_int32), ("tile_cells", C.c_int32), ("max_tile_particles", C.c_int32),
                ("loose_particles", C.c_int32), ("det_rebins", C.c_int64), ("deterministic", C.c_int32),
                ("pad_", C.c_int32)]


class Tuning(C.Structure):
    """mpmw_tuning: the launch configuration in use and the autotuner's
    state and medians."""
    _fields_ = [("state", C.c_int32), ("tile_order", C.c_int32), ("grid_order", C.c_int32),
                ("grid_block", C.c_int32), ("post_order", C.c_int32), ("trials", C.c_int32),
                ("tile_ms", C.c_float * 2), ("grid_ms", C.c_float * 6), ("post_ms", C.c_float * 2)]

    def as_dict(self):
        return {"state": self.state, "tile_order": self.tile_order, "grid_order": self.grid_order,
                "grid_block": self.grid_block, "post_order": self.post_order, "trials": self.trials,
                "tile_ms": list(self.tile_ms), "grid_ms": list(self.grid_ms), "post_ms": list(self.post_ms)}


def launch_config(tile_order=0, grid_order=1, grid_block=256, post_order=1):
    t = Tuning()
    t.tile_order, t.grid_order, t.grid_block, t.post_order = tile_order, grid_order, grid_block, post_order
    return t


class MPMError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libmpmw_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MPMError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "mpmw_last_error": (C.c_char_p, []),
        "mpmw_abi_version": (C.c_int, []),
        "mpmw_params_default": (None, [C.POINTER(Params)]),
        "mpmw_create": (C.c_int, [C.POINTER(Params), C.POINTER(vp)]),
        "mpmw_destroy": (C.c_int, [vp]),
        "mpmw_reset": (C.c_int, [vp]),
        "mpmw_set_particles": (C.c_int, [vp, C.POINTER(ParticleArrays)]),
        "mpmw_get_particles": (C.c_int, [vp, C.POINTER(ParticleArrays)]),
        "mpmw_particle_count": (C.c_int, [vp, _ip]),
        "mpmw_init_particle_box": (C.c_int, [vp, _dp, _dp, C.c_double]),
        "mpmw_init_cylinder": (C.c_int, [vp, _dp, C.c_double, C.c_double, C.c_double, C.c_int32]),
        "mpmw_gen_jittered_slab": (C.c_int, [C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_double, _dp,
                                             C.c_double, C.c_double, _dp, _dp, _dp, _ip, _ip]),
        "mpmw_init_jittered_slab": (C.c_int, [vp, C.c_uint64, C.c_int32, C.c_int32, C.c_int32, _dp, _dp,
                                              C.c_double]),
        "mpmw_set_tool_samples": (C.c_int, [vp, _dp, C.c_int32, C.c_double]),
        "mpmw_set_tool_analytic": (C.c_int, [vp, C.POINTER(ToolDesc)]),
        "mpmw_set_tool_pose": (C.c_int, [vp, C.POINTER(ToolPose)]),
        "mpmw_set_tool_sdf_mesh": (C.c_int, [vp, C.POINTER(ToolDesc), _dp, C.c_int32, _ip, _dp, C.c_int32,
                                             C.c_double, C.c_int32]),
        "mpmw_get_tool_sdf": (C.c_int, [vp, C.POINTER(C.c_float), _ip, _dp, _dp]),
        "mpmw_set_transfer": (C.c_int, [vp, C.c_int32]),
        "mpmw_get_affine": (C.c_int, [vp, _dp, C.c_int32]),
        "mpmw_set_affine": (C.c_int, [vp, _dp, C.c_int32]),
        "mpmw_export_ndjson": (C.c_int, [vp, C.c_int32, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "mpmw_export_vtk": (C.c_int, [vp, C.c_char_p]),
        "mpmw_export_ndjson_noalloc": (C.c_int, [vp, C.c_int32, C.c_char_p, C.c_size_t,
                                                 C.POINTER(C.c_size_t)]),
        "mpmw_morph_tool_sdf": (C.c_int, [vp, _dp, C.c_int32]),
        "mpmw_tool_mesh_params_default": (None, [C.POINTER(ToolMeshParams)]),
        "mpmw_parametric_tool_mesh": (C.c_int, [C.POINTER(ToolMeshParams), _dp, _ip, _dp, _ip, _ip]),
        "mpmw_set_tool_tip": (C.c_int, [vp, C.c_double, C.c_double, C.c_double]),
        "mpmw_apply_external_force": (C.c_int, [vp, _dp, C.c_double, _dp]),
        "mpmw_step": (C.c_int, [vp, C.c_double]),
        "mpmw_run": (C.c_int, [vp, C.c_int32, C.c_double]),
        "mpmw_synchronize": (C.c_int, [vp]),
        "mpmw_stable_dt": (C.c_int, [vp, _dp]),
        "mpmw_get_metrics": (C.c_int, [vp, C.POINTER(Metrics)]),
        "mpmw_sync_metrics": (C.c_int, [vp]),
        "mpmw_compact": (C.c_int, [vp]),
        "mpmw_get_bounds": (C.c_int, [vp, _dp]),
        "mpmw_temperature_at": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, _dp]),
        "mpmw_apply_heat_flux": (C.c_int, [vp, C.c_double, C.c_double, C.c_double, C.c_double]),
        "mpmw_set_particle_temperatures": (C.c_int, [vp, _dp, C.c_int32]),
        "mpmw_set_deterministic": (C.c_int, [vp, C.c_int32]),
        "mpmw_get_grid_info": (C.c_int, [vp, C.POINTER(GridInfo)]),
        "mpmw_set_grid_snapshot": (C.c_int, [vp, C.c_int32]),
        "mpmw_get_grid_snapshot": (C.c_int, [vp, _dp, _dp, _dp, _dp]),
        "mpmw_set_kernel_timing": (C.c_int, [vp, C.c_int32]),
        "mpmw_get_kernel_times": (C.c_int, [vp, C.POINTER(KernelTimes)]),
        "mpmw_get_stream": (C.c_int, [vp, C.POINTER(vp)]),
        "mpmw_device_bytes": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "mpmw_get_diagnostics": (C.c_int, [vp, C.POINTER(Diagnostics)]),
        "mpmw_batch_create": (C.c_int, [C.c_int32, C.POINTER(vp)]),
        "mpmw_batch_destroy": (C.c_int, [vp]),
        "mpmw_batch_add": (C.c_int, [vp, vp]),
        "mpmw_batch_remove": (C.c_int, [vp, vp]),
        "mpmw_batch_count": (C.c_int, [vp, _ip]),
        "mpmw_batch_step": (C.c_int, [vp, _dp, C.c_int32]),
        "mpmw_batch_synchronize": (C.c_int, [vp]),
        "mpmw_set_graph": (C.c_int, [vp, C.c_int32]),
        "mpmw_graph_stats": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "mpmw_batch_set_graph": (C.c_int, [vp, C.c_int32]),
        "mpmw_batch_graph_stats": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "mpmw_set_autotune": (C.c_int, [vp, C.c_int32]),
        "mpmw_set_launch_config": (C.c_int, [vp, C.POINTER(Tuning)]),
        "mpmw_get_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
        "mpmw_batch_set_autotune": (C.c_int, [vp, C.c_int32]),
        "mpmw_batch_set_launch_config": (C.c_int, [vp, C.POINTER(Tuning)]),
        "mpmw_batch_get_tuning": (C.c_int, [vp, C.POINTER(Tuning)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        msg = lib().mpmw_last_error().decode(errors="replace")
        raise MPMError(f"{what} failed ({rc}): {msg}")


def _p(a, ct=_dp):
    return None if a is None else a.ctypes.data_as(ct)


def default_params(**kw):
    p = Params()
    lib().mpmw_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def make_arrays(n):
    return {
        "pos": np.zeros((n, 3)), "vel": np.zeros((n, 3)), "fext": np.zeros((n, 3)),
        "mass": np.zeros(n), "volume": np.zeros(n), "density": np.zeros(n),
        "temperature": np.zeros(n), "plastic_strain": np.zeros(n), "strain_rate": np.zeros(n),
        "damage": np.zeros(n), "residual_stress": np.zeros(n),
        "stress": np.zeros((n, 6)), "backstress": np.zeros((n, 6)),
        "id": np.zeros(n, dtype=np.int32), "status": np.zeros(n, dtype=np.int32),
    }


def _arrays_struct(d):
    a = ParticleArrays()
    a.n = len(d["mass"])
    keep = {}
    for k in ("pos", "vel", "fext", "mass", "volume", "density", "temperature", "plastic_strain",
              "strain_rate", "damage", "residual_stress", "stress", "backstress"):
        v = d.get(k)
        if v is not None:
            v = np.ascontiguousarray(v, dtype=np.float64)
            keep[k] = v
        setattr(a, k, _p(v))
    for k in ("id", "status"):
        v = np.ascontiguousarray(d[k], dtype=np.int32)
        keep[k] = v
        setattr(a, k, _p(v, _ip))
    return a, keep


def gen_jittered_slab(seed, cx, cy, cz, h, origin=(0.0, 0.0, 0.0), density=4430.0, fixture=0.0):
    n = cx * cy * cz * 8
    d = make_arrays(n)
    d["density"][:] = density
    org = np.asarray(origin, dtype=np.float64)
    _check(lib().mpmw_gen_jittered_slab(seed, cx, cy, cz, h, _p(org), density, fixture, _p(d["pos"]),
                                        _p(d["mass"]), _p(d["volume"]), _p(d["id"], _ip),
                                        _p(d["status"], _ip)), "gen_jittered_slab")
    return d


class Solver:
    """One MPM workpiece on one MI355X (a handle on the C ABI)."""

    def __init__(self, params=None, **kw):
        self.params = params if params is not None else default_params(**kw)
        h = C.c_void_p()
        _check(lib().mpmw_create(C.byref(self.params), C.byref(h)), "mpmw_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mpmw_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_particles(self, d):
        a, keep = _arrays_struct(d)
        _check(lib().mpmw_set_particles(self.h, C.byref(a)), "set_particles")

    def count(self):
        n = C.c_int32()
        _check(lib().mpmw_particle_count(self.h, C.byref(n)), "particle_count")
        return n.value

    def get_particles(self):
        d = make_arrays(self.count())
        a, keep = _arrays_struct(d)
        _check(lib().mpmw_get_particles(self.h, C.byref(a)), "get_particles")
        for k, v in keep.items():
            d[k] = v
        return d

    def init_box(self, mn, mx, spacing):
        _check(lib().mpmw_init_particle_box(self.h, _p(np.asarray(mn, float)), _p(np.asarray(mx, float)),
                                            spacing), "init_particle_box")

    def init_cylinder(self, c, radius, length, spacing, axis=2):
        _check(lib().mpmw_init_cylinder(self.h, _p(np.asarray(c, float)), radius, length, spacing, axis),
               "init_cylinder")

    def init_jittered_slab(self, seed, cx, cy, cz, origin=(0, 0, 0), vel=(0, 0, 0), temperature=293.0):
        _check(lib().mpmw_init_jittered_slab(self.h, seed, cx, cy, cz, _p(np.asarray(origin, float)),
                                             _p(np.asarray(vel, float)), temperature), "init_jittered_slab")

    def set_tool_samples(self, xyz, contact_radius):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        _check(lib().mpmw_set_tool_samples(self.h, _p(xyz), len(xyz), contact_radius), "set_tool_samples")

    def set_tool_analytic(self, desc):
        _check(lib().mpmw_set_tool_analytic(self.h, C.byref(desc)), "set_tool_analytic")

    def set_tool_sdf_mesh(self, desc, verts, tris, normals, padding, voxels_per_axis):
        """Voxel SDF tool from a triangle mesh (rest frame) with the contact
        law of `desc` (band, friction, grad_eps)."""
        v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        n = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
        if len(n) != len(t):
            raise MPMError("set_tool_sdf_mesh: one normal per triangle")
        _check(lib().mpmw_set_tool_sdf_mesh(self.h, C.byref(desc), _p(v), len(v),
                                            t.ctypes.data_as(_ip), _p(n), len(t), padding, voxels_per_axis),
               "set_tool_sdf_mesh")

    def get_tool_sdf(self):
        """(grid[nz, ny, nx] float32, origin, voxel) of the voxel SDF in use."""
        dims = (C.c_int32 * 3)()
        org = (C.c_double * 3)()
        vox = C.c_double()
        _check(lib().mpmw_get_tool_sdf(self.h, None, dims, org, C.byref(vox)), "get_tool_sdf")
        nx, ny, nz = dims
        g = np.empty((nz, ny, nx), dtype=np.float32)
        _check(lib().mpmw_get_tool_sdf(self.h, g.ctypes.data_as(C.POINTER(C.c_float)), dims, org, C.byref(vox)),
               "get_tool_sdf")
        return g, tuple(org), vox.value

    def set_transfer(self, mode):
        """TRANSFER_PIC (trilinear), TRANSFER_APIC (quadratic; every C set to zero) or
        TRANSFER_APIC_TEXTBOOK (quadratic with C = 4/h^2 B and no 0.99 decay)."""
        _check(lib().mpmw_set_transfer(self.h, mode), "set_transfer")

    def get_affine(self):
        n = self.count()
        C9 = np.zeros((n, 9))
        _check(lib().mpmw_get_affine(self.h, _p(C9), n), "get_affine")
        return C9

    def set_affine(self, C9):
        C9 = np.ascontiguousarray(C9, dtype=np.float64).reshape(-1, 9)
        _check(lib().mpmw_set_affine(self.h, _p(C9), len(C9)), "set_affine")

    def morph_tool_sdf(self, wear):
        """Move the voxel tool's mesh vertices inward by wear[i] along their
        vertex normals and rebuild its SDF."""
        w = np.ascontiguousarray(wear, dtype=np.float64).reshape(-1)
        _check(lib().mpmw_morph_tool_sdf(self.h, _p(w), len(w)), "morph_tool_sdf")

    def export_ndjson(self, step):
        """The NDJSON particle message of the current state (bytes, newline
        included; b"" when every particle is INACTIVE)."""
        n = C.c_size_t(0)
        _check(lib().mpmw_export_ndjson(self.h, int(step), None, 0, C.byref(n)), "export_ndjson")
        if n.value == 0:
            return b""
        buf = np.empty(n.value, dtype=np.uint8)  # filled by the call, no zeroing
        _check(lib().mpmw_export_ndjson(self.h, int(step), buf.ctypes.data_as(C.c_char_p), n.value, C.byref(n)),
               "export_ndjson")
        return buf[:n.value].tobytes()

    def export_ndjson_noalloc(self, step):
        """The same message as export_ndjson, built in the library's static
        per-device scratch array (no device or pinned allocation; at most
        about 2.5 M particles)."""
        n = C.c_size_t(0)
        _check(lib().mpmw_export_ndjson_noalloc(self.h, int(step), None, 0, C.byref(n)), "export_ndjson_noalloc")
        if n.value == 0:
            return b""
        buf = np.empty(n.value, dtype=np.uint8)
        _check(lib().mpmw_export_ndjson_noalloc(self.h, int(step), buf.ctypes.data_as(C.c_char_p), n.value,
                                                C.byref(n)), "export_ndjson_noalloc")
        return buf[:n.value].tobytes()

    def export_vtk(self, path):
        """Write the legacy VTK particle file of the current state."""
        _check(lib().mpmw_export_vtk(self.h, os.fsencode(path)), "export_vtk")

    def set_tool_pose(self, pose):
        _check(lib().mpmw_set_tool_pose(self.h, C.byref(pose)), "set_tool_pose")

    def set_tool_tip(self, x, y, z):
        _check(lib().mpmw_set_tool_tip(self.h, x, y, z), "set_tool_tip")

    def apply_external_force(self, c, radius, f):
        _check(lib().mpmw_apply_external_force(self.h, _p(np.asarray(c, float)), radius,
                                               _p(np.asarray(f, float))), "apply_external_force")

    def step(self, dt):
        _check(lib().mpmw_step(self.h, dt), "step")

    def run(self, nsteps, dt):
        _check(lib().mpmw_run(self.h, nsteps, dt), "run")

    def synchronize(self):
        _check(lib().mpmw_synchronize(self.h), "synchronize")

    def stable_dt(self):
        v = C.c_double()
        _check(lib().mpmw_stable_dt(self.h, C.byref(v)), "stable_dt")
        return v.value

    def metrics(self):
        m = Metrics()
        _check(lib().mpmw_get_metrics(self.h, C.byref(m)), "get_metrics")
        return m

    def sync_metrics(self):
        _check(lib().mpmw_sync_metrics(self.h), "sync_metrics")

    def compact(self):
        _check(lib().mpmw_compact(self.h), "compact")

    def reset(self):
        _check(lib().mpmw_reset(self.h), "reset")

    def bounds(self):
        b = np.zeros(6)
        _check(lib().mpmw_get_bounds(self.h, _p(b)), "get_bounds")
        return b

    def temperature_at(self, x, y, z):
        v = C.c_double()
        _check(lib().mpmw_temperature_at(self.h, x, y, z, C.byref(v)), "temperature_at")
        return v.value

    def apply_heat_flux(self, x, y, z, q):
        _check(lib().mpmw_apply_heat_flux(self.h, x, y, z, q), "apply_heat_flux")

    def set_particle_temperatures(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        _check(lib().mpmw_set_particle_temperatures(self.h, _p(t), len(t)), "set_particle_temperatures")

    def set_graph(self, on):
        """Replay the step's launches as a HIP graph (captured anew when the
        launch configuration changes)."""
        _check(lib().mpmw_set_graph(self.h, 1 if on else 0), "set_graph")

    def graph_stats(self):
        v = (C.c_int64 * 2)()
        _check(lib().mpmw_graph_stats(self.h, v), "graph_stats")
        return {"captures": v[0], "replays": v[1]}

    def set_autotune(self, trials):
        """Tune the step's launch configuration over the next steps (trials
        steps per candidate; 0 stops); results are unchanged by it."""
        _check(lib().mpmw_set_autotune(self.h, trials), "set_autotune")

    def set_launch_config(self, cfg):
        _check(lib().mpmw_set_launch_config(self.h, C.byref(cfg)), "set_launch_config")

    def tuning(self):
        t = Tuning()
        _check(lib().mpmw_get_tuning(self.h, C.byref(t)), "get_tuning")
        return t.as_dict()

    def set_deterministic(self, on):
        _check(lib().mpmw_set_deterministic(self.h, 1 if on else 0), "set_deterministic")

    def grid_info(self):
        g = GridInfo()
        _check(lib().mpmw_get_grid_info(self.h, C.byref(g)), "get_grid_info")
        return g

    def set_grid_snapshot(self, on):
        _check(lib().mpmw_set_grid_snapshot(self.h, 1 if on else 0), "set_grid_snapshot")

    def grid_snapshot(self):
        g = self.grid_info()
        n = g.nx * g.ny * g.nz
        m, p, f, tm = np.zeros(n), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
        _check(lib().mpmw_get_grid_snapshot(self.h, _p(m), _p(p), _p(f), _p(tm)), "get_grid_snapshot")
        return {"mass": m, "mom": p, "force": f, "tm": tm}

    def set_kernel_timing(self, on):
        """True / TIMING_ALL: every kernel class; TIMING_PARTICLE_PASS: the
        fused particle pass only; False / TIMING_OFF: none."""
        mode = TIMING_ALL if on is True else (TIMING_OFF if on is False else int(on))
        _check(lib().mpmw_set_kernel_timing(self.h, mode), "set_kernel_timing")

    def kernel_times(self):
        t = KernelTimes()
        _check(lib().mpmw_get_kernel_times(self.h, C.byref(t)), "get_kernel_times")
        return list(t.ms), list(t.launches)

    def stream(self):
        v = C.c_void_p()
        _check(lib().mpmw_get_stream(self.h, C.byref(v)), "get_stream")
        return v.value

    def diagnostics(self):
        d = Diagnostics()
        _check(lib().mpmw_get_diagnostics(self.h, C.byref(d)), "get_diagnostics")
        return {k: getattr(d, k) for k, _ in Diagnostics._fields_ if k != "pad_"}

    def device_bytes(self):
        v = C.c_int64()
        _check(lib().mpmw_device_bytes(self.h, C.byref(v)), "device_bytes")
        return v.value


class Batch:
    """Solvers stepped together: each kernel of the step runs once for all
    members over a flattened block list (mpmw_batch_*).  The batch holds
    references to its members; removing or closing does not destroy them."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().mpmw_batch_create(device, C.byref(h)), "mpmw_batch_create")
        self.h = h
        self.members = []

    def close(self):
        if getattr(self, "h", None):
            lib().mpmw_batch_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, solver):
        _check(lib().mpmw_batch_add(self.h, solver.h), "batch_add")
        self.members.append(solver)

    def remove(self, solver):
        _check(lib().mpmw_batch_remove(self.h, solver.h), "batch_remove")
        self.members.remove(solver)

    def count(self):
        self.members = [m for m in self.members if getattr(m, "h", None)]
        n = C.c_int32()
        _check(lib().mpmw_batch_count(self.h, C.byref(n)), "batch_count")
        return n.value

    def step(self, dt):
        """One step of every member; dt is a scalar or one value per member."""
        self.members = [m for m in self.members if getattr(m, "h", None)]  # closed solvers left the batch
        n = len(self.members)
        d = np.full(n, float(dt)) if np.isscalar(dt) else np.ascontiguousarray(dt, dtype=np.float64)
        _check(lib().mpmw_batch_step(self.h, _p(d), n), "batch_step")

    def synchronize(self):
        _check(lib().mpmw_batch_synchronize(self.h), "batch_synchronize")

    def set_graph(self, on):
        _check(lib().mpmw_batch_set_graph(self.h, 1 if on else 0), "batch_set_graph")

    def graph_stats(self):
        v = (C.c_int64 * 2)()
        _check(lib().mpmw_batch_graph_stats(self.h, v), "batch_graph_stats")
        return {"captures": v[0], "replays": v[1]}

    def set_autotune(self, trials):
        _check(lib().mpmw_batch_set_autotune(self.h, trials), "batch_set_autotune")

    def set_launch_config(self, cfg):
        _check(lib().mpmw_batch_set_launch_config(self.h, C.byref(cfg)), "batch_set_launch_config")

    def tuning(self):
        t = Tuning()
        _check(lib().mpmw_batch_get_tuning(self.h, C.byref(t)), "batch_get_tuning")
        return t.as_dict()
