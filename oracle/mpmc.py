"""ctypes binding of the CPU FP64 restatement of the reference's tool FEM
step (FEMSolver::step, src/FEMSolver.cu) and of its spatial-hash XPBD
contact (ContactSolver::resolveContacts, src/ContactSolver.cu;
SharedSpatialHash.cu), in mpmc_oracle.c (built into liborc_mpms.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it as the checker; the product library never does.
"""
import ctypes as C

import numpy as np

import mpms

OK, WORN, FAILED, FIXED = 0, 1, 2, 3
REGION_SHANK = 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class ToolParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("young", "poisson", "density", "thermal_conductivity", "specific_heat",
                                          "damping_ratio", "mass_scaling", "coolant_htc", "spindle_stiffness",
                                          "spindle_damping")] + [
        ("spindle_enabled", C.c_int32), ("thermal_skip", C.c_int32), ("polar_textbook", C.c_int32),
        ("pad_", C.c_int32)]


class Node(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x", "y", "z", "ox", "oy", "oz", "vx", "vy", "vz", "fx", "fy", "fz",
                                          "mass", "temperature", "von_mises", "wear")] + [
        ("s", C.c_double * 6), ("stress_weight", C.c_double)] + [
        (n, C.c_double) for n in ("heat", "conduction", "fn", "pen")] + [("loff", C.c_double * 3)] + [
        (n, C.c_int32) for n in ("id", "status", "fixed", "in_contact", "driven", "region")]


NODE_DTYPE = np.dtype({"names": [f[0] for f in Node._fields_],
                       "formats": [np.float64 if t is C.c_double else
                                   (np.int32 if t is C.c_int32 else (np.float64, t._length_))
                                   for _, t in Node._fields_],
                       "offsets": [getattr(Node, f[0]).offset for f in Node._fields_],
                       "itemsize": C.sizeof(Node)})

CONTACT_DOUBLES = ("contact_radius", "contact_stiffness", "contact_damping_ratio", "friction", "heat_partition",
                   "workpiece_shear_yield", "max_contact_pressure", "jc_A", "jc_B", "jc_n", "jc_C", "jc_m",
                   "jc_rate_ref", "jc_T_ref", "jc_T_melt", "max_penetration_fraction", "tool_specific_heat",
                   "work_specific_heat", "work_physical_density", "tool_melt_temp", "work_melt_temp", "contact_htc",
                   "emissivity", "stefan_boltzmann", "ambient_temp", "tool_physical_density", "latent_heat_fusion",
                   "solidus_temp", "liquidus_temp")


class ContactParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in CONTACT_DOUBLES] + [
        ("jc_enabled", C.c_int32), ("engagement_ramp_steps", C.c_int32)]


_done = False


def lib():
    global _done
    L = mpms.lib()
    if _done:
        return L
    vp = C.c_void_p
    sig = {
        "omt_params_default": (None, [C.POINTER(ToolParams)]),
        "omt_create": (vp, [C.POINTER(ToolParams)]),
        "omt_destroy": (None, [vp]),
        "omt_set_mesh": (None, [vp, _dp, _ip, C.c_int, _dp, _ip, C.c_int, _ip, C.c_int, _ip, _ip, _dp, C.c_int]),
        "omt_init_driven": (None, [vp, C.c_double]),
        "omt_set_spindle_state": (None, [vp, _dp, _dp]),
        "omt_set_transform": (None, [vp, _dp, _dp, C.c_double]),
        "omt_set_rigid_velocity": (None, [vp, _dp, C.c_double, C.c_double, C.c_double]),
        "omt_apply_contact_force": (None, [vp, C.c_int, _dp, C.c_double]),
        "omt_step": (None, [vp, C.c_double]),
        "omt_counts": (C.c_int, [vp, C.POINTER(C.c_int)]),
        "omt_get_surface": (None, [vp, vp]),
        "omt_get_tets": (None, [vp, vp]),
        "omt_set_surface": (None, [vp, vp]),
        "omt_set_tets": (None, [vp, vp]),
        "omt_get_elements": (None, [vp, _ip, _dp, _dp]),
        "omt_polar": (None, [_dp, C.c_int, _dp]),
        "omc_params_default": (None, [C.POINTER(ContactParams)]),
        "omc_create": (vp, [C.POINTER(ContactParams)]),
        "omc_destroy": (None, [vp]),
        "omc_resolve": (None, [vp, vp, vp, C.c_double]),
        "omc_results": (None, [vp, _dp]),
        "omc_hash_size": (C.c_int, [vp]),
        "omc_hash": (None, [vp, _ip, _ip, _ip, C.c_int]),
        "omc_node_hash": (C.c_int32, [C.c_double, C.c_double, C.c_double, _dp, C.c_double, C.c_int32]),
        "omc_next_prime": (C.c_int32, [C.c_int32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _done = True
    return L


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def tool_params(**kw):
    p = ToolParams()
    lib().omt_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def contact_params(**kw):
    p = ContactParams()
    lib().omc_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class ToolOracle:
    """FEMSolver's data and step (dynamic mode) restated."""

    def __init__(self, params=None, **kw):
        self.params = params if params is not None else tool_params(**kw)
        self.h = lib().omt_create(C.byref(self.params))

    def __del__(self):
        if getattr(self, "h", None):
            lib().omt_destroy(self.h)
            self.h = None

    def set_mesh(self, mesh):
        """mesh: dict with surf_xyz (ns,3), surf_region (ns), tet_xyz (nt,3),
        tet_region (nt), elems (ne,4), emb_surf (nemb), emb_tet (nemb,4),
        emb_w (nemb,4)."""
        s, t = _d(mesh["surf_xyz"]).reshape(-1, 3), _d(mesh["tet_xyz"]).reshape(-1, 3)
        sr, tr = _i(mesh["surf_region"]), _i(mesh["tet_region"])
        el = _i(mesh["elems"]).reshape(-1, 4)
        es, et, ew = _i(mesh["emb_surf"]), _i(mesh["emb_tet"]).reshape(-1, 4), _d(mesh["emb_w"]).reshape(-1, 4)
        self._keep = (s, t, sr, tr, el, es, et, ew)
        L = lib()
        L.omt_set_mesh(self.h, s.ctypes.data_as(_dp), sr.ctypes.data_as(_ip), len(s), t.ctypes.data_as(_dp),
                       tr.ctypes.data_as(_ip), len(t), el.ctypes.data_as(_ip), len(el), es.ctypes.data_as(_ip),
                       et.ctypes.data_as(_ip), ew.ctypes.data_as(_dp), len(es))

    def init_driven(self, top_fraction):
        lib().omt_init_driven(self.h, top_fraction)

    def set_spindle_state(self, pos, vel):
        lib().omt_set_spindle_state(self.h, _d(pos).ctypes.data_as(_dp), _d(vel).ctypes.data_as(_dp))

    def set_transform(self, pos, axis, omega):
        lib().omt_set_transform(self.h, _d(pos).ctypes.data_as(_dp), _d(axis).ctypes.data_as(_dp), omega)

    def set_rigid_velocity(self, v, omega_z, cx, cy):
        lib().omt_set_rigid_velocity(self.h, _d(v).ctypes.data_as(_dp), omega_z, cx, cy)

    def apply_contact_force(self, node, f, heat):
        lib().omt_apply_contact_force(self.h, node, _d(f).ctypes.data_as(_dp), heat)

    def step(self, dt):
        lib().omt_step(self.h, dt)

    def counts(self):
        c = (C.c_int * 4)()
        lib().omt_counts(self.h, c)
        return tuple(c)

    def surface(self):
        a = np.zeros(self.counts()[0], dtype=NODE_DTYPE)
        lib().omt_get_surface(self.h, a.ctypes.data_as(C.c_void_p))
        return a

    def tets(self):
        a = np.zeros(self.counts()[1], dtype=NODE_DTYPE)
        lib().omt_get_tets(self.h, a.ctypes.data_as(C.c_void_p))
        return a

    def set_surface(self, a):
        a = np.ascontiguousarray(a, dtype=NODE_DTYPE)
        assert len(a) == self.counts()[0]
        lib().omt_set_surface(self.h, a.ctypes.data_as(C.c_void_p))

    def set_tets(self, a):
        a = np.ascontiguousarray(a, dtype=NODE_DTYPE)
        assert len(a) == self.counts()[1]
        lib().omt_set_tets(self.h, a.ctypes.data_as(C.c_void_p))

    def elements(self):
        ne = self.counts()[2]
        n4, v, m = np.zeros((ne, 4), np.int32), np.zeros(ne), np.zeros((ne, 9))
        lib().omt_get_elements(self.h, n4.ctypes.data_as(_ip), v.ctypes.data_as(_dp), m.ctypes.data_as(_dp))
        return n4, v, m.reshape(ne, 3, 3)


class ContactOracle:
    """ContactSolver's hash path restated, acting on an mpms.Oracle's
    particles and a ToolOracle's surface nodes."""

    def __init__(self, params=None, **kw):
        self.params = params if params is not None else contact_params(**kw)
        self.h = lib().omc_create(C.byref(self.params))

    def __del__(self):
        if getattr(self, "h", None):
            lib().omc_destroy(self.h)
            self.h = None

    def resolve(self, sph, tool, dt):
        lib().omc_resolve(self.h, sph.h, tool.h, dt)

    def results(self):
        r = np.zeros(5)
        lib().omc_results(self.h, r.ctypes.data_as(_dp))
        return {"contacts": int(r[0]), "heat": r[1], "force": r[2], "cumulative_heat": r[3], "engagement": r[4]}

    def hash_table(self, nn):
        T = lib().omc_hash_size(self.h)
        s, e, o = np.zeros(T, np.int32), np.zeros(T, np.int32), np.zeros(nn, np.int32)
        lib().omc_hash(self.h, s.ctypes.data_as(_ip), e.ctypes.data_as(_ip), o.ctypes.data_as(_ip), nn)
        return s, e, o


def polar(F, textbook=False):
    R = np.zeros(9)
    lib().omt_polar(_d(F).reshape(9).ctypes.data_as(_dp), 1 if textbook else 0, R.ctypes.data_as(_dp))
    return R.reshape(3, 3)


def node_hash(p, origin, cell, table):
    return lib().omc_node_hash(p[0], p[1], p[2], _d(origin).ctypes.data_as(_dp), cell, table)


def next_prime(n):
    return lib().omc_next_prime(n)
