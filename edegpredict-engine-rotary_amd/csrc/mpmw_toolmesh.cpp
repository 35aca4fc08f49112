// mpmw_toolmesh.cpp -- the parametric tool surface mesh of the reference's
// application (generateParametricToolMesh, main.cpp:83-233), for the voxel
// SDF tool when no mesh file is given.  Host code, no device work.
//
// A cylinder of radius r from z = 0 down to z = -L as 21 rings of 32 nodes
// (ring row at z = -row / 20 * L, node i at angle 2 pi i / 32), two
// triangles per quad, a top cap (a cone apex at z = h for the drill and the
// thread mill, else a centre node at z = 0) and a bottom centre node at
// z = -L; 674 nodes and 1344 triangles.  Each triangle's normal is the
// normalised cross product (v1 - v0) x (v2 - v0) (Vec3::normalized,
// Types.h:183-186: zero when the length is not above 1e-12).
#include <cmath>
#include <cstdint>

#include "mpmw.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kCirc = 32, kAxial = 20;

struct MeshOut {
    double *v;
    int32_t *t;
    double *n;
    int32_t nv = 0, nt = 0;

    int node(double x, double y, double z) {
        v[3 * nv] = x;
        v[3 * nv + 1] = y;
        v[3 * nv + 2] = z;
        return nv++;
    }
    void tri(int a, int b, int c) {
#pragma clang fp contract(off)
        t[3 * nt] = a;
        t[3 * nt + 1] = b;
        t[3 * nt + 2] = c;
        const double *p0 = v + 3 * a, *p1 = v + 3 * b, *p2 = v + 3 * c;
        const double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        double q[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double len = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        for (int k = 0; k < 3; ++k) n[3 * nt + k] = len > 1e-12 ? q[k] / len : 0.0;
        ++nt;
    }
};

}  // namespace

extern "C" {

void mpmw_tool_mesh_params_default(mpmw_tool_mesh_params *p) {
    if (!p) return;
    p->kind = MPMW_TOOLMESH_DRILL;
    p->boring_bar_length_mm = 40.0;
    p->boring_bar_diameter_mm = 10.0;
    p->reamer_diameter_mm = 10.0;
    p->thread_major_diameter_mm = 10.0;
    p->thread_form_angle_deg = 60.0;
}

int mpmw_parametric_tool_mesh(const mpmw_tool_mesh_params *p, double *verts, int32_t *tris, double *normals,
                              int32_t *nverts, int32_t *ntris) {
#pragma clang fp contract(off)
    if (!p || !verts || !tris || !normals) return MPMW_E_INVAL;
    double radius, length, cone_h = 0.0;
    bool cone = false;
    switch (p->kind) {  // main.cpp:115-176
    case MPMW_TOOLMESH_DRILL:
        radius = 0.005;
        length = 0.025;
        cone = true;
        cone_h = radius / std::tan(59.0 * kPi / 180.0);
        break;
    case MPMW_TOOLMESH_BORING:
        radius = (p->boring_bar_diameter_mm / 2.0) / 1000.0;
        length = p->boring_bar_length_mm / 1000.0;
        break;
    case MPMW_TOOLMESH_REAMING:
        radius = (p->reamer_diameter_mm / 2.0) / 1000.0;
        length = 0.020;
        break;
    case MPMW_TOOLMESH_THREADING:
        radius = (p->thread_major_diameter_mm / 2.0 * 0.8) / 1000.0;
        length = 0.015;
        cone = true;
        cone_h = radius / std::tan(p->thread_form_angle_deg / 2.0 * kPi / 180.0);
        break;
    case MPMW_TOOLMESH_MILLING:
        radius = 0.005;
        length = 0.025;
        break;
    default:
        return MPMW_E_INVAL;
    }
    if (!(radius > 0.0) || !(length > 0.0) || !std::isfinite(radius) || !std::isfinite(length) ||
        !std::isfinite(cone_h))
        return MPMW_E_INVAL;
    MeshOut m{verts, tris, normals};
    // cylindrical body (:182-202)
    const int body = m.nv;
    for (int row = 0; row <= kAxial; ++row) {
        const double z = -(double)row / kAxial * length;
        for (int i = 0; i < kCirc; ++i) {
            const double th = 2.0 * kPi * i / kCirc;
            m.node(radius * std::cos(th), radius * std::sin(th), z);
        }
    }
    for (int row = 0; row < kAxial; ++row)
        for (int i = 0; i < kCirc; ++i) {
            const int i2 = (i + 1) % kCirc;
            const int a = body + row * kCirc + i, b = body + row * kCirc + i2;
            const int c = body + (row + 1) * kCirc + i, d = body + (row + 1) * kCirc + i2;
            m.tri(a, b, d);
            m.tri(a, d, c);
        }
    // top cap (:204-217)
    if (cone) {
        const int apex = m.node(0.0, 0.0, cone_h);
        for (int i = 0; i < kCirc; ++i) m.tri(apex, body + i, body + (i + 1) % kCirc);
    } else {
        const int centre = m.node(0.0, 0.0, 0.0);
        for (int i = 0; i < kCirc; ++i) m.tri(centre, body + (i + 1) % kCirc, body + i);
    }
    // bottom cap (:219-225)
    const int ring = body + kAxial * kCirc;
    const int bottom = m.node(0.0, 0.0, -length);
    for (int i = 0; i < kCirc; ++i) m.tri(bottom, ring + i, ring + (i + 1) % kCirc);
    if (nverts) *nverts = m.nv;
    if (ntris) *ntris = m.nt;
    return MPMW_OK;
}

}  // extern "C"
