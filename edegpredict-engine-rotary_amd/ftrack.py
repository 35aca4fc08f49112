"""ctypes binding of the MI355X F-tracking MPM solver (include/mpms.h, in
libmpmw_hip.so): the reference's MPMSolver step (src/MPMSolver.cu) with FP32
bulk and FP64 islands, or FP64 throughout.  No CPU fallback: without the HIP
library or a device every call fails."""
import ctypes as C

import numpy as np

import mpmw

FP32, FP64 = 0, 1
LOD_ACTIVE, LOD_NEAR, LOD_FAR = 0, 1, 2
ELASTIC_NEOHOOKEAN, ELASTIC_COROTATED = 0, 1
RETURN_RADIAL, RETURN_NEWTON = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

PARAM_DOUBLES = ("young", "poisson", "density", "mass_scaling", "specific_heat", "taylor_quinney", "melt_temp",
                 "ambient_temp", "jc_A", "jc_B", "jc_n", "jc_C", "jc_m", "kinematic_hardening", "jc_D1", "jc_D2",
                 "jc_D3", "jc_D4", "jc_D5", "damage_threshold", "ref_strain_rate", "lod_active_radius",
                 "lod_near_radius", "fixture_thickness", "fixture_damping", "dx")


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in PARAM_DOUBLES] + [
        ("domain_min", C.c_double * 3), ("domain_max", C.c_double * 3), ("cfl", C.c_double),
        ("lod_enabled", C.c_int32), ("max_particles", C.c_int32), ("elastic_model", C.c_int32),
        ("return_map", C.c_int32), ("precision", C.c_int32), ("device", C.c_int32),
        ("island_det", C.c_double), ("newton_fp64", C.c_int32), ("polar_fp64", C.c_int32)]


class Arrays(C.Structure):
    _fields_ = [("n", C.c_int32), ("pad_", C.c_int32)] + [(k, _dp) for k in (
        "pos", "vel", "fext", "mass", "volume", "temperature", "F", "Fp", "C", "stress", "backstress",
        "plastic_strain", "strain_rate", "damage", "residual_stress")] + [
        ("id", _ip), ("status", _ip), ("lod", _ip)]


class Diagnostics(C.Structure):
    _fields_ = [("box_lo", C.c_int32 * 3), ("box_n", C.c_int32 * 3), ("fits", C.c_int64), ("sorts", C.c_int64),
                ("overflow", C.c_int64), ("island_kinematics", C.c_int64), ("island_newton", C.c_int64),
                ("out_of_tile", C.c_int64)]


SHAPES = {"pos": 3, "vel": 3, "fext": 3, "mass": 1, "volume": 1, "temperature": 1, "F": 9, "Fp": 9, "C": 9,
          "stress": 6, "backstress": 6, "plastic_strain": 1, "strain_rate": 1, "damage": 1, "residual_stress": 1}

_done = False


def lib():
    global _done
    L = mpmw.lib()
    if _done:
        return L
    vp = C.c_void_p
    sig = {
        "mpms_params_default": (None, [C.POINTER(Params)]),
        "mpms_create": (C.c_int, [C.POINTER(Params), C.POINTER(vp)]),
        "mpms_destroy": (C.c_int, [vp]),
        "mpms_set_particles": (C.c_int, [vp, C.POINTER(Arrays)]),
        "mpms_get_particles": (C.c_int, [vp, C.POINTER(Arrays)]),
        "mpms_particle_count": (C.c_int, [vp, _ip]),
        "mpms_init_particle_box": (C.c_int, [vp, _dp, _dp, C.c_double]),
        "mpms_init_cylinder": (C.c_int, [vp, _dp, C.c_double, C.c_double, C.c_double]),
        "mpms_set_tool_nodes": (C.c_int, [vp, _dp, _dp, C.c_int32]),
        "mpms_set_tool_position": (C.c_int, [vp, C.c_double, C.c_double, C.c_double]),
        "mpms_apply_external_force": (C.c_int, [vp, _dp, C.c_double, _dp]),
        "mpms_step": (C.c_int, [vp, C.c_double]),
        "mpms_run": (C.c_int, [vp, C.c_int32, C.c_double]),
        "mpms_synchronize": (C.c_int, [vp]),
        "mpms_stable_dt": (C.c_int, [vp, _dp]),
        "mpms_metrics": (C.c_int, [vp, _dp]),
        "mpms_get_bounds": (C.c_int, [vp, _dp]),
        "mpms_get_grid": (C.c_int, [vp, _ip, _ip, _dp, _dp, _dp]),
        "mpms_get_diagnostics": (C.c_int, [vp, C.POINTER(Diagnostics)]),
        "mpms_get_stream": (C.c_int, [vp, C.POINTER(vp)]),
        "mpms_set_graph": (C.c_int, [vp, C.c_int32]),
        "mpms_graph_stats": (C.c_int, [vp, C.POINTER(C.c_int64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _done = True
    return L


EXPORTS = ("mpms_params_default", "mpms_create", "mpms_destroy", "mpms_set_particles", "mpms_get_particles",
           "mpms_particle_count", "mpms_init_particle_box", "mpms_init_cylinder", "mpms_set_tool_nodes",
           "mpms_set_tool_position", "mpms_apply_external_force", "mpms_step", "mpms_run", "mpms_synchronize",
           "mpms_stable_dt", "mpms_metrics", "mpms_get_bounds", "mpms_get_grid", "mpms_get_diagnostics",
           "mpms_get_stream", "mpms_set_graph", "mpms_graph_stats")


def _check(rc, what):
    if rc != 0:
        raise mpmw.MPMError(f"{what} failed ({rc}): {lib().mpmw_last_error().decode(errors='replace')}")


def default_params(**kw):
    p = Params()
    lib().mpms_params_default(C.byref(p))
    for k, v in kw.items():
        if k in ("domain_min", "domain_max"):
            getattr(p, k)[:] = v
        else:
            setattr(p, k, v)
    return p


def make_arrays(n):
    d = {k: np.zeros((n, c) if c > 1 else n) for k, c in SHAPES.items()}
    for k in ("id", "status", "lod"):
        d[k] = np.zeros(n, dtype=np.int32)
    return d


def _struct(d):
    a = Arrays()
    a.n = len(d["mass"]) if "mass" in d else len(d["id"])
    keep = {}
    for k in SHAPES:
        v = d.get(k)
        if v is None:
            setattr(a, k, None)
            continue
        v = np.ascontiguousarray(v, dtype=np.float64)
        keep[k] = v
        setattr(a, k, v.ctypes.data_as(_dp))
    for k in ("id", "status", "lod"):
        v = d.get(k)
        if v is None:
            setattr(a, k, None)
            continue
        v = np.ascontiguousarray(v, dtype=np.int32)
        keep[k] = v
        setattr(a, k, v.ctypes.data_as(_ip))
    return a, keep


class Solver:
    def __init__(self, params=None, **kw):
        self.params = params if params is not None else default_params(**kw)
        h = C.c_void_p()
        _check(lib().mpms_create(C.byref(self.params), C.byref(h)), "mpms_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mpms_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_particles(self, d):
        a, keep = _struct(d)
        _check(lib().mpms_set_particles(self.h, C.byref(a)), "mpms_set_particles")
        del keep

    def count(self):
        n = C.c_int32()
        _check(lib().mpms_particle_count(self.h, C.byref(n)), "mpms_particle_count")
        return n.value

    def get_particles(self, fields=None):
        """Every field, or only those named (id and status always): a
        16 M-particle state is ~7 GB of FP64 arrays."""
        d = make_arrays(self.count())
        if fields is not None:
            want = set(fields) | {"id", "status"}
            d = {k: v for k, v in d.items() if k in want}
        a, keep = _struct(d)
        _check(lib().mpms_get_particles(self.h, C.byref(a)), "mpms_get_particles")
        for k, v in keep.items():
            d[k] = v
        return d

    def init_box(self, mn, mx, spacing):
        _check(lib().mpms_init_particle_box(self.h, np.asarray(mn, float).ctypes.data_as(_dp),
                                            np.asarray(mx, float).ctypes.data_as(_dp), spacing), "init_box")

    def init_cylinder(self, c, radius, length, spacing):
        _check(lib().mpms_init_cylinder(self.h, np.asarray(c, float).ctypes.data_as(_dp), radius, length, spacing),
               "init_cylinder")

    def set_tool_nodes(self, xyz, vel):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        _check(lib().mpms_set_tool_nodes(self.h, xyz.ctypes.data_as(_dp), vel.ctypes.data_as(_dp), len(xyz)),
               "set_tool_nodes")

    def set_tool_position(self, x, y, z):
        _check(lib().mpms_set_tool_position(self.h, x, y, z), "set_tool_position")

    def apply_external_force(self, c, radius, f):
        _check(lib().mpms_apply_external_force(self.h, np.asarray(c, float).ctypes.data_as(_dp), radius,
                                               np.asarray(f, float).ctypes.data_as(_dp)), "apply_external_force")

    def step(self, dt):
        _check(lib().mpms_step(self.h, dt), "mpms_step")

    def run(self, n, dt):
        _check(lib().mpms_run(self.h, n, dt), "mpms_run")

    def synchronize(self):
        _check(lib().mpms_synchronize(self.h), "mpms_synchronize")

    def stable_dt(self):
        v = C.c_double()
        _check(lib().mpms_stable_dt(self.h, C.byref(v)), "stable_dt")
        return v.value

    def metrics(self):
        m = np.zeros(4)
        _check(lib().mpms_metrics(self.h, m.ctypes.data_as(_dp)), "metrics")
        return {"max_stress": m[0], "max_temperature": m[1], "kinetic_energy": m[2], "internal_energy": m[3]}

    def bounds(self):
        b = np.zeros(6)
        _check(lib().mpms_get_bounds(self.h, b.ctypes.data_as(_dp)), "bounds")
        return b

    def grid(self):
        lo, n = (C.c_int32 * 3)(), (C.c_int32 * 3)()
        _check(lib().mpms_get_grid(self.h, lo, n, None, None, None), "get_grid")
        nn = n[0] * n[1] * n[2]
        m, p, f = np.zeros(nn), np.zeros((nn, 3)), np.zeros((nn, 3))
        _check(lib().mpms_get_grid(self.h, lo, n, m.ctypes.data_as(_dp), p.ctypes.data_as(_dp),
                                   f.ctypes.data_as(_dp)), "get_grid")
        return {"lo": tuple(lo), "n": tuple(n), "mass": m, "mom": p, "force": f}

    def diagnostics(self):
        d = Diagnostics()
        _check(lib().mpms_get_diagnostics(self.h, C.byref(d)), "diagnostics")
        return {"box_lo": tuple(d.box_lo), "box_n": tuple(d.box_n), "fits": d.fits, "sorts": d.sorts,
                "overflow": d.overflow, "island_kinematics": d.island_kinematics, "island_newton": d.island_newton,
                "out_of_tile": d.out_of_tile}

    def set_graph(self, on=True):
        _check(lib().mpms_set_graph(self.h, 1 if on else 0), "mpms_set_graph")

    def graph_stats(self):
        out = (C.c_int64 * 2)()
        _check(lib().mpms_graph_stats(self.h, out), "mpms_graph_stats")
        return {"captures": out[0], "replays": out[1]}

    def stream(self):
        v = C.c_void_p()
        _check(lib().mpms_get_stream(self.h, C.byref(v)), "get_stream")
        return v.value
