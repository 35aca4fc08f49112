"""CPU restatement of the reference application's parametric tool mesh
(generateParametricToolMesh, src/main.cpp:83-233); test infrastructure only
(tests/ import it, the product library never does).

Tool kinds and dimensions (:115-176): drill r = 5 mm, L = 25 mm, cone tip of
height r / tan(59 deg); boring bar r = diameter / 2, L = bar length (config
defaults 10 mm and 40 mm); reamer r = diameter / 2 (10 mm), L = 20 mm, flat;
thread mill r = 0.8 x major diameter / 2, L = 15 mm, tip height
r / tan(form angle / 2); end mill r = 5 mm, L = 25 mm, flat.  Mesh
(:178-225): 21 rings of 32 nodes, ring `row` at z = -row / 20 * L, node i at
angle 2 pi i / 32; quads as (a, b, d), (a, d, c); a cone apex or flat centre
on top, a centre at z = -L below.  Normals: (v1 - v0) x (v2 - v0) divided by
its length when that exceeds 1e-12 (Vec3::normalized, include/Types.h), else
zero.
"""
import math

import numpy as np

DRILL, BORING, REAMING, THREADING, MILLING = range(5)
N_CIRC, N_AXIAL = 32, 20


def dimensions(kind, boring_bar_length_mm=40.0, boring_bar_diameter_mm=10.0, reamer_diameter_mm=10.0,
               thread_major_diameter_mm=10.0, thread_form_angle_deg=60.0):
    if kind == DRILL:
        r = 0.005
        return r, 0.025, r / math.tan(59.0 * math.pi / 180.0)
    if kind == BORING:
        return (boring_bar_diameter_mm / 2.0) / 1000.0, boring_bar_length_mm / 1000.0, None
    if kind == REAMING:
        return (reamer_diameter_mm / 2.0) / 1000.0, 0.020, None
    if kind == THREADING:
        r = (thread_major_diameter_mm / 2.0 * 0.8) / 1000.0
        return r, 0.015, r / math.tan(thread_form_angle_deg / 2.0 * math.pi / 180.0)
    if kind == MILLING:
        return 0.005, 0.025, None
    raise ValueError(kind)


def tool_mesh(kind=DRILL, **params):
    r, length, cone_h = dimensions(kind, **params)
    V, T, N = [], [], []

    def node(x, y, z):
        V.append((x, y, z))
        return len(V) - 1

    def tri(a, b, c):
        p0, p1, p2 = (np.array(V[i], dtype=np.float64) for i in (a, b, c))
        e1, e2 = p1 - p0, p2 - p0
        q = [e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]]
        ln = math.sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2])
        T.append((a, b, c))
        N.append(tuple(c_ / ln for c_ in q) if ln > 1e-12 else (0.0, 0.0, 0.0))

    for row in range(N_AXIAL + 1):
        z = -float(row) / N_AXIAL * length
        for i in range(N_CIRC):
            th = 2.0 * math.pi * i / N_CIRC
            node(r * math.cos(th), r * math.sin(th), z)
    for row in range(N_AXIAL):
        for i in range(N_CIRC):
            i2 = (i + 1) % N_CIRC
            a, b = row * N_CIRC + i, row * N_CIRC + i2
            c, d = (row + 1) * N_CIRC + i, (row + 1) * N_CIRC + i2
            tri(a, b, d)
            tri(a, d, c)
    if cone_h is not None:
        apex = node(0.0, 0.0, cone_h)
        for i in range(N_CIRC):
            tri(apex, i, (i + 1) % N_CIRC)
    else:
        centre = node(0.0, 0.0, 0.0)
        for i in range(N_CIRC):
            tri(centre, (i + 1) % N_CIRC, i)
    ring = N_AXIAL * N_CIRC
    bottom = node(0.0, 0.0, -length)
    for i in range(N_CIRC):
        tri(bottom, ring + i, ring + (i + 1) % N_CIRC)
    return np.array(V), np.array(T, dtype=np.int32), np.array(N)
