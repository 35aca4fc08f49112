"""ctypes binding of the CPU FP64 restatement of the reference's F-tracking
MPM step, MPMSolver::step (src/MPMSolver.cu), in mpms_oracle.c.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it as the checker; the product library never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liborc_mpms.so")

ACTIVE, INACTIVE, CHIP, BOUNDARY, FIXED = 0, 1, 2, 3, 4
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
        ("return_map", C.c_int32)]


class Particle(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x", "y", "z", "vx", "vy", "vz", "efx", "efy", "efz", "mass", "volume",
                                          "density", "temperature")] + [
        ("F", C.c_double * 9), ("Fp", C.c_double * 9), ("C", C.c_double * 9), ("s", C.c_double * 6),
        ("a", C.c_double * 6)] + [(n, C.c_double) for n in ("plastic_strain", "strain_rate", "damage",
                                                             "residual_stress")] + [
        ("id", C.c_int32), ("status", C.c_int32), ("lod", C.c_int32), ("pad_", C.c_int32)]


# structured dtype with the same layout, for bulk copies
PARTICLE_DTYPE = np.dtype({"names": [f[0] for f in Particle._fields_],
                           "formats": [np.float64 if t is C.c_double else
                                       (np.int32 if t is C.c_int32 else (np.float64, t._length_))
                                       for _, t in Particle._fields_],
                           "offsets": [getattr(Particle, f[0]).offset for f in Particle._fields_],
                           "itemsize": C.sizeof(Particle)})

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"F-tracking oracle not built: {LIB_PATH} (run make -C oracle)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "oms_params_default": (None, [C.POINTER(Params)]),
        "oms_create": (vp, [C.POINTER(Params)]),
        "oms_destroy": (None, [vp]),
        "oms_set_particles": (None, [vp, vp, C.c_int]),
        "oms_particle_count": (C.c_int, [vp]),
        "oms_get_particles": (None, [vp, vp]),
        "oms_set_tool_nodes": (None, [vp, _dp, _dp, C.c_int]),
        "oms_set_tool_position": (None, [vp, C.c_double, C.c_double, C.c_double]),
        "oms_apply_external_force": (None, [vp, _dp, C.c_double, _dp]),
        "oms_step": (None, [vp, C.c_double]),
        "oms_stable_dt": (C.c_double, [vp]),
        "oms_metrics": (None, [vp, _dp]),
        "oms_bounds": (None, [vp, _dp]),
        "oms_init_box": (None, [vp, _dp, _dp, C.c_double]),
        "oms_init_cylinder": (None, [vp, _dp, C.c_double, C.c_double, C.c_double]),
        "oms_grid": (None, [vp, _ip, _ip, _dp, _dp, _dp]),
        "oms_polar": (None, [_dp, _dp, C.POINTER(C.c_int)]),
        "oms_p2g_stress": (None, [C.POINTER(Params), C.POINTER(Particle), _dp, _dp, C.POINTER(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(_dp)


def default_params(**kw):
    p = Params()
    lib().oms_params_default(C.byref(p))
    for k, v in kw.items():
        if k in ("domain_min", "domain_max"):
            getattr(p, k)[:] = v
        else:
            setattr(p, k, v)
    return p


def new_particles(n):
    """n particles with MPMParticle's default values (Types.h:246-265)."""
    a = np.zeros(n, dtype=PARTICLE_DTYPE)
    a["temperature"] = 25.0
    a["id"] = -1
    eye = np.eye(3).reshape(9)
    a["F"] = eye
    a["Fp"] = eye
    return a


class Oracle:
    def __init__(self, params=None, **kw):
        self.params = params if params is not None else default_params(**kw)
        self.h = lib().oms_create(C.byref(self.params))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oms_destroy(self.h)
            self.h = None

    def set_particles(self, a):
        a = np.ascontiguousarray(a, dtype=PARTICLE_DTYPE)
        lib().oms_set_particles(self.h, a.ctypes.data_as(C.c_void_p), len(a))

    def count(self):
        return lib().oms_particle_count(self.h)

    def get_particles(self):
        a = np.zeros(self.count(), dtype=PARTICLE_DTYPE)
        lib().oms_get_particles(self.h, a.ctypes.data_as(C.c_void_p))
        return a

    def set_tool_nodes(self, xyz, vel):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        lib().oms_set_tool_nodes(self.h, _p(xyz), _p(vel), len(xyz))

    def set_tool_position(self, x, y, z):
        lib().oms_set_tool_position(self.h, x, y, z)

    def apply_external_force(self, c, radius, f):
        lib().oms_apply_external_force(self.h, _p(np.asarray(c, float)), radius, _p(np.asarray(f, float)))

    def step(self, dt):
        lib().oms_step(self.h, dt)

    def run(self, n, dt):
        for _ in range(n):
            self.step(dt)

    def stable_dt(self):
        return lib().oms_stable_dt(self.h)

    def metrics(self):
        m = np.zeros(4)
        lib().oms_metrics(self.h, _p(m))
        return {"max_stress": m[0], "max_temperature": m[1], "kinetic_energy": m[2], "internal_energy": m[3]}

    def bounds(self):
        b = np.zeros(6)
        lib().oms_bounds(self.h, _p(b))
        return b

    def init_box(self, mn, mx, spacing):
        lib().oms_init_box(self.h, _p(np.asarray(mn, float)), _p(np.asarray(mx, float)), spacing)

    def init_cylinder(self, c, radius, length, spacing):
        lib().oms_init_cylinder(self.h, _p(np.asarray(c, float)), radius, length, spacing)

    def grid(self):
        """The last step's node sub-box: lo, n and mass, momentum (after the
        grid update), force (of the P2G) per node (x fastest)."""
        lo, n = (C.c_int32 * 3)(), (C.c_int32 * 3)()
        lib().oms_grid(self.h, lo, n, None, None, None)
        nn = n[0] * n[1] * n[2]
        m, p, f = np.zeros(nn), np.zeros((nn, 3)), np.zeros((nn, 3))
        lib().oms_grid(self.h, lo, n, _p(m), _p(p), _p(f))
        return {"lo": tuple(lo), "n": tuple(n), "mass": m, "mom": p, "force": f}


def polar(F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    R = np.zeros(9)
    it = C.c_int()
    lib().oms_polar(_p(F), _p(R), C.byref(it))
    return R.reshape(3, 3), it.value


def p2g_stress(params, particle):
    """The constitutive part of the P2G on one particle (a structured scalar
    of PARTICLE_DTYPE, updated in place): (sigma 3x3, J, skip)."""
    a = np.ascontiguousarray(np.atleast_1d(particle), dtype=PARTICLE_DTYPE)
    sig = np.zeros(9)
    J = C.c_double()
    skip = C.c_int()
    lib().oms_p2g_stress(C.byref(params), C.cast(a.ctypes.data, C.POINTER(Particle)), _p(sig), C.byref(J),
                         C.byref(skip))
    return sig.reshape(3, 3), J.value, skip.value, a[0]
