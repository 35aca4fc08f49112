"""ctypes binding of the MI355X tool FEM step and the spatial-hash XPBD
contact (include/mpmc.h, in libmpmw_hip.so): the reference's FEMSolver in its
dynamic mode (src/FEMSolver.cu) and ContactSolver's hash path
(src/ContactSolver.cu), coupled to the F-tracking solver of ftrack.py.  No
CPU fallback: without the HIP library or a device every call fails.

cylinder_tool_mesh builds a tool of the shape the benchmark uses (a
tetrahedralised cylinder with its boundary nodes as the surface nodes); the
reference takes its tool mesh from its geometry loader, which is not part of
this library.
"""
import ctypes as C

import numpy as np

import ftrack
import mpmw

OK, WORN, FAILED = 0, 1, 2
REGION_TOOL_BODY, REGION_CUTTING_EDGE, REGION_SHANK = 1, 2, 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class ToolParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("young", "poisson", "density", "thermal_conductivity", "specific_heat",
                                          "damping_ratio", "mass_scaling", "coolant_htc", "spindle_stiffness",
                                          "spindle_damping")] + [
        ("spindle_enabled", C.c_int32), ("thermal_skip", C.c_int32), ("polar_textbook", C.c_int32),
        ("device", C.c_int32)]


class Mesh(C.Structure):
    _fields_ = [("n_surface", C.c_int32), ("n_tet_nodes", C.c_int32), ("n_elements", C.c_int32),
                ("n_embeds", C.c_int32), ("surface_xyz", _dp), ("surface_region", _ip), ("tet_xyz", _dp),
                ("tet_region", _ip), ("elements", _ip), ("embed_surface", _ip), ("embed_tets", _ip),
                ("embed_weights", _dp)]


class NodeArrays(C.Structure):
    _fields_ = [(n, _dp) for n in ("pos", "vel", "force", "mass", "temperature", "von_mises", "wear", "stress")] + [
        ("status", _ip), ("in_contact", _ip)]


CONTACT_DOUBLES = ("contact_radius", "contact_stiffness", "contact_damping_ratio", "friction", "heat_partition",
                   "workpiece_shear_yield", "max_contact_pressure", "jc_A", "jc_B", "jc_n", "jc_C", "jc_m",
                   "jc_rate_ref", "jc_T_ref", "jc_T_melt", "max_penetration_fraction", "tool_specific_heat",
                   "work_specific_heat", "work_physical_density", "tool_melt_temp", "work_melt_temp", "contact_htc",
                   "emissivity", "stefan_boltzmann", "ambient_temp", "tool_physical_density", "latent_heat_fusion",
                   "solidus_temp", "liquidus_temp")


class ContactParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in CONTACT_DOUBLES] + [
        ("jc_enabled", C.c_int32), ("engagement_ramp_steps", C.c_int32), ("device", C.c_int32), ("pad_", C.c_int32)]


class Results(C.Structure):
    _fields_ = [("contacts", C.c_int64), ("heat", C.c_double), ("force", C.c_double),
                ("cumulative_heat", C.c_double), ("engagement", C.c_double), ("probed", C.c_int64),
                ("table_size", C.c_int32), ("pad_", C.c_int32)]


EXPORTS = ("mpmt_params_default", "mpmt_create", "mpmt_destroy", "mpmt_set_mesh", "mpmt_init_driven_nodes",
           "mpmt_set_spindle_state", "mpmt_set_transform", "mpmt_set_rigid_velocity", "mpmt_apply_contact_force",
           "mpmt_step", "mpmt_synchronize", "mpmt_counts", "mpmt_get_surface", "mpmt_get_tet_nodes",
           "mpmt_set_surface", "mpmc_params_default", "mpmc_create", "mpmc_destroy", "mpmc_set_params",
           "mpmc_resolve", "mpmc_get_results", "mpmc_link_tool_nodes")

_done = False


def lib():
    global _done
    L = ftrack.lib()
    if _done:
        return L
    vp = C.c_void_p
    sig = {
        "mpmt_params_default": (None, [C.POINTER(ToolParams)]),
        "mpmt_create": (C.c_int, [C.POINTER(ToolParams), C.POINTER(vp)]),
        "mpmt_destroy": (C.c_int, [vp]),
        "mpmt_set_mesh": (C.c_int, [vp, C.POINTER(Mesh)]),
        "mpmt_init_driven_nodes": (C.c_int, [vp, C.c_double]),
        "mpmt_set_spindle_state": (C.c_int, [vp, _dp, _dp]),
        "mpmt_set_transform": (C.c_int, [vp, _dp, _dp, C.c_double]),
        "mpmt_set_rigid_velocity": (C.c_int, [vp, _dp, C.c_double, C.c_double, C.c_double]),
        "mpmt_apply_contact_force": (C.c_int, [vp, C.c_int32, _dp, C.c_double]),
        "mpmt_step": (C.c_int, [vp, C.c_double]),
        "mpmt_synchronize": (C.c_int, [vp]),
        "mpmt_counts": (C.c_int, [vp, _ip]),
        "mpmt_get_surface": (C.c_int, [vp, C.POINTER(NodeArrays)]),
        "mpmt_get_tet_nodes": (C.c_int, [vp, C.POINTER(NodeArrays)]),
        "mpmt_set_surface": (C.c_int, [vp, _dp, _dp]),
        "mpmc_params_default": (None, [C.POINTER(ContactParams)]),
        "mpmc_create": (C.c_int, [C.POINTER(ContactParams), C.POINTER(vp)]),
        "mpmc_destroy": (C.c_int, [vp]),
        "mpmc_set_params": (C.c_int, [vp, C.POINTER(ContactParams)]),
        "mpmc_resolve": (C.c_int, [vp, vp, vp, C.c_double]),
        "mpmc_get_results": (C.c_int, [vp, C.POINTER(Results)]),
        "mpmc_link_tool_nodes": (C.c_int, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _done = True
    return L


def _check(rc, what):
    if rc != 0:
        raise mpmw.MPMError(f"{what} failed ({rc}): {lib().mpmw_last_error().decode(errors='replace')}")


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def tool_params(**kw):
    p = ToolParams()
    lib().mpmt_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def contact_params(**kw):
    p = ContactParams()
    lib().mpmc_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def cylinder_tool_mesh(radius, length, h, base=(0.0, 0.0, 0.0), shank_fraction=0.2, face_samples=0):
    """A cylindrical tool along +z from `base` (its tip face at base z):
    grid points at spacing h inside the radius, each grid cube whose 8
    corners are inside cut into 6 tets around its main diagonal (a conforming
    Kuhn subdivision), tet nodes = the points the tets use, surface nodes =
    the tet nodes on boundary faces (a face of one tet only), each tied to
    its tet node with weight 1.  With face_samples = n > 2 every boundary
    triangle also carries the interior points of its barycentric grid of
    order n ((n - 1)(n - 2) / 2 points) as surface nodes, embedded in the
    triangle's three tet nodes with those barycentric weights: a surface
    finer than the tets, as the reference's tool meshes are.  Tet nodes in
    the top shank_fraction of the length are ToolRegion SHANK, the bottom
    10 % CUTTING_EDGE, the rest TOOL_BODY."""
    nr = int(np.floor(radius / h))
    nz = int(round(length / h))
    ii = np.arange(-nr, nr + 1)
    gx, gy, gz = np.meshgrid(ii, ii, np.arange(nz + 1), indexing="ij")
    inside = (gx * h) ** 2 + (gy * h) ** 2 <= radius * radius * (1 + 1e-12)
    # cubes with all 8 corners inside
    c = inside[:-1, :-1, :-1] & inside[1:, :-1, :-1] & inside[:-1, 1:, :-1] & inside[1:, 1:, :-1] & \
        inside[:-1, :-1, 1:] & inside[1:, :-1, 1:] & inside[:-1, 1:, 1:] & inside[1:, 1:, 1:]
    ci, cj, ck = np.nonzero(c)
    dims = np.array(gx.shape)

    def gid(i, j, k):
        return (i * dims[1] + j) * dims[2] + k

    corners = {}
    for bits in range(8):
        di, dj, dk = bits & 1, (bits >> 1) & 1, (bits >> 2) & 1
        corners[bits] = gid(ci + di, cj + dj, ck + dk)
    tets = []
    for a, b in ((1, 2), (2, 1), (1, 4), (4, 1), (2, 4), (4, 2)):
        # path 0 -> e_a -> e_a + e_b -> 7 through the cube
        tets.append(np.stack([corners[0], corners[a], corners[a | b], corners[7]], axis=1))
    el = np.concatenate(tets, axis=0)
    used, inv = np.unique(el, return_inverse=True)
    el = inv.reshape(el.shape).astype(np.int32)
    ui, uj, uk = used // (dims[1] * dims[2]), (used // dims[2]) % dims[1], used % dims[2]
    xyz = np.stack([base[0] + (ui - nr) * h, base[1] + (uj - nr) * h, base[2] + uk * h], axis=1)
    # boundary faces
    faces = np.concatenate([el[:, [0, 1, 2]], el[:, [0, 1, 3]], el[:, [0, 2, 3]], el[:, [1, 2, 3]]], axis=0)
    faces = np.sort(faces, axis=1)
    uf, cnt = np.unique(faces, axis=0, return_counts=True)
    surf = np.unique(uf[cnt == 1].reshape(-1))
    zrel = (xyz[:, 2] - base[2]) / max(length, 1e-30)
    region = np.where(zrel >= 1.0 - shank_fraction, REGION_SHANK,
                      np.where(zrel <= 0.1, REGION_CUTTING_EDGE, REGION_TOOL_BODY)).astype(np.int32)
    emb_tet = np.full((len(surf), 4), -1, np.int32)
    emb_tet[:, 0] = surf
    emb_w = np.zeros((len(surf), 4))
    emb_w[:, 0] = 1.0
    sxyz, sreg = xyz[surf], region[surf]
    if face_samples > 2:
        n = face_samples
        bary = np.array([(i, j, n - i - j) for i in range(1, n) for j in range(1, n - i)], float) / n
        tri = uf[cnt == 1]                                   # (nf, 3) tet-node indices
        w = np.repeat(bary[None], len(tri), 0).reshape(-1, 3)
        nodes = np.repeat(tri, len(bary), 0)
        pts = (xyz[nodes] * w[:, :, None]).sum(1)
        e_t = np.full((len(pts), 4), -1, np.int32)
        e_t[:, :3] = nodes
        e_w = np.zeros((len(pts), 4))
        e_w[:, :3] = w
        emb_tet, emb_w = np.concatenate([emb_tet, e_t]), np.concatenate([emb_w, e_w])
        sxyz = np.concatenate([sxyz, pts])
        sreg = np.concatenate([sreg, region[nodes[:, 0]]])
    return {"surf_xyz": sxyz, "surf_region": sreg.astype(np.int32), "tet_xyz": xyz, "tet_region": region,
            "elems": el, "emb_surf": np.arange(len(sxyz), dtype=np.int32), "emb_tet": emb_tet, "emb_w": emb_w}


class Tool:
    """FEMSolver on the GPU: mesh, spindle, transform, step, node read-back."""

    def __init__(self, params=None, **kw):
        self.params = params if params is not None else tool_params(**kw)
        h = C.c_void_p()
        _check(lib().mpmt_create(C.byref(self.params), C.byref(h)), "mpmt_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mpmt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mesh(self, mesh):
        s, t = _d(mesh["surf_xyz"]).reshape(-1, 3), _d(mesh["tet_xyz"]).reshape(-1, 3)
        sr, tr = _i(mesh["surf_region"]), _i(mesh["tet_region"])
        el = _i(mesh["elems"]).reshape(-1, 4)
        es, et, ew = _i(mesh["emb_surf"]), _i(mesh["emb_tet"]).reshape(-1, 4), _d(mesh["emb_w"]).reshape(-1, 4)
        m = Mesh(len(s), len(t), len(el), len(es), s.ctypes.data_as(_dp), sr.ctypes.data_as(_ip),
                 t.ctypes.data_as(_dp), tr.ctypes.data_as(_ip), el.ctypes.data_as(_ip), es.ctypes.data_as(_ip),
                 et.ctypes.data_as(_ip), ew.ctypes.data_as(_dp))
        _check(lib().mpmt_set_mesh(self.h, C.byref(m)), "mpmt_set_mesh")

    def init_driven(self, top_fraction):
        _check(lib().mpmt_init_driven_nodes(self.h, top_fraction), "mpmt_init_driven_nodes")

    def set_spindle_state(self, pos, vel):
        _check(lib().mpmt_set_spindle_state(self.h, _d(pos).ctypes.data_as(_dp), _d(vel).ctypes.data_as(_dp)),
               "mpmt_set_spindle_state")

    def set_transform(self, pos, axis, omega):
        _check(lib().mpmt_set_transform(self.h, _d(pos).ctypes.data_as(_dp), _d(axis).ctypes.data_as(_dp), omega),
               "mpmt_set_transform")

    def set_rigid_velocity(self, v, omega_z, cx, cy):
        _check(lib().mpmt_set_rigid_velocity(self.h, _d(v).ctypes.data_as(_dp), omega_z, cx, cy),
               "mpmt_set_rigid_velocity")

    def apply_contact_force(self, node, f, heat):
        _check(lib().mpmt_apply_contact_force(self.h, node, _d(f).ctypes.data_as(_dp), heat),
               "mpmt_apply_contact_force")

    def step(self, dt):
        _check(lib().mpmt_step(self.h, dt), "mpmt_step")

    def synchronize(self):
        _check(lib().mpmt_synchronize(self.h), "mpmt_synchronize")

    def counts(self):
        c = (C.c_int32 * 4)()
        _check(lib().mpmt_counts(self.h, c), "mpmt_counts")
        return tuple(c)

    def _nodes(self, fn, n):
        d = {"pos": np.zeros((n, 3)), "vel": np.zeros((n, 3)), "force": np.zeros((n, 3)), "mass": np.zeros(n),
             "temperature": np.zeros(n), "von_mises": np.zeros(n), "wear": np.zeros(n), "stress": np.zeros((n, 6)),
             "status": np.zeros(n, np.int32), "in_contact": np.zeros(n, np.int32)}
        a = NodeArrays(*[d[k].ctypes.data_as(_dp) for k in ("pos", "vel", "force", "mass", "temperature",
                                                             "von_mises", "wear", "stress")],
                       d["status"].ctypes.data_as(_ip), d["in_contact"].ctypes.data_as(_ip))
        _check(fn(self.h, C.byref(a)), "node read-back")
        return d

    def surface(self):
        return self._nodes(lib().mpmt_get_surface, self.counts()[0])

    def tets(self):
        return self._nodes(lib().mpmt_get_tet_nodes, self.counts()[1])

    def set_surface(self, temperature=None, force=None):
        t = None if temperature is None else _d(temperature)
        f = None if force is None else _d(force).reshape(-1, 3)
        _check(lib().mpmt_set_surface(self.h, None if t is None else t.ctypes.data_as(_dp),
                                      None if f is None else f.ctypes.data_as(_dp)), "mpmt_set_surface")


class Contact:
    """ContactSolver's hash path on the GPU between an ftrack.Solver's
    particles and a Tool's surface nodes."""

    def __init__(self, params=None, **kw):
        self.params = params if params is not None else contact_params(**kw)
        h = C.c_void_p()
        _check(lib().mpmc_create(C.byref(self.params), C.byref(h)), "mpmc_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mpmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        self.params = params
        _check(lib().mpmc_set_params(self.h, C.byref(params)), "mpmc_set_params")

    def resolve(self, solver, tool, dt):
        _check(lib().mpmc_resolve(self.h, solver.h, tool.h, dt), "mpmc_resolve")

    def results(self):
        r = Results()
        _check(lib().mpmc_get_results(self.h, C.byref(r)), "mpmc_get_results")
        return {"contacts": r.contacts, "heat": r.heat, "force": r.force, "cumulative_heat": r.cumulative_heat,
                "engagement": r.engagement, "probed": r.probed, "table_size": r.table_size}


def link_tool_nodes(solver, tool):
    """The F-tracking solver's rasterised tool = the tool's surface nodes."""
    _check(lib().mpmc_link_tool_nodes(solver.h, tool.h), "mpmc_link_tool_nodes")
