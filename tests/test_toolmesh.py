"""The parametric tool mesh (generateParametricToolMesh, main.cpp:83-233).

The restatement (oracle/toolmesh.py) is pinned by closed forms, derived by
hand from the reference's construction: 674 nodes and 1344 triangles; ring
nodes at radius r on 21 planes z = -row L / 20; the drill's apex at
r / tan 59 deg and the thread mill's at r / tan(form / 2); a closed surface
(every edge in exactly two triangles) with unit normals; and the signed
volume of the reference's winding, -P L for the flat-topped tools (body and
caps wound inward) and -P (L - h / 3) for the pointed ones (the cone cap is
wound outward), P the area of the 32-gon of radius r.  The library's host
function equals the restatement bit for bit for every tool kind and for
overridden machining parameters.  No device is needed.
"""
import math
from collections import Counter

import numpy as np
import pytest

import mpmw
import toolmesh

KINDS = [(toolmesh.DRILL, {}), (toolmesh.BORING, {}), (toolmesh.REAMING, {}), (toolmesh.THREADING, {}),
         (toolmesh.MILLING, {}), (toolmesh.BORING, {"boring_bar_length_mm": 55.0, "boring_bar_diameter_mm": 8.0}),
         (toolmesh.THREADING, {"thread_major_diameter_mm": 12.0, "thread_form_angle_deg": 55.0})]


@pytest.mark.parametrize("kind,over", KINDS)
def test_restatement_closed_forms(kind, over):
    V, T, N = toolmesh.tool_mesh(kind, **over)
    r, L, h = toolmesh.dimensions(kind, **over)
    assert V.shape == (674, 3) and T.shape == (1344, 3)
    ring = V[:21 * 32]
    assert np.allclose(np.hypot(ring[:, 0], ring[:, 1]), r, rtol=1e-15, atol=0)
    assert np.allclose(ring[:, 2].reshape(21, 32), (-np.arange(21) / 20.0 * L)[:, None], rtol=0, atol=1e-18)
    if kind == toolmesh.DRILL:
        assert V[672, 2] == pytest.approx(0.005 / math.tan(math.radians(59.0)), rel=1e-15)
    assert V[673].tolist() == [0.0, 0.0, -L]
    edges = Counter(tuple(sorted((int(t[a]), int(t[b])))) for t in T for a, b in ((0, 1), (1, 2), (2, 0)))
    assert set(edges.values()) == {2}
    assert np.allclose(np.linalg.norm(N, axis=1), 1.0, rtol=0, atol=1e-15)
    vol = sum(float(np.dot(V[a], np.cross(V[b], V[c]))) for a, b, c in T) / 6.0
    P = 16.0 * r * r * math.sin(2.0 * math.pi / 32)
    want = -P * (L - h / 3.0) if h is not None else -P * L
    assert vol == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("kind,over", KINDS)
def test_library_matches_restatement(kind, over):
    V, T, N = mpmw.parametric_tool_mesh(kind, **over)
    RV, RT, RN = toolmesh.tool_mesh(kind, **over)
    assert np.array_equal(T, RT)
    assert np.array_equal(V, RV)
    assert np.array_equal(N, RN)


def test_library_rejects_unknown_kind():
    with pytest.raises(mpmw.MPMError):
        mpmw.parametric_tool_mesh(9)
