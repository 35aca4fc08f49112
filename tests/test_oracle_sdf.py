"""Pins of the oracle's voxel SDF (ToolSDF restated, oracle/mpm_oracle.c).

* point-triangle distance: vertex and edge regions against the exact
  distance (projection on the plane, segments) computed independently here;
  the interior region against the reference's own formula for the point it
  takes as closest, a + u (b - a) + v (c - a) with u the barycentric weight
  of a and v that of b (ToolSDF.cu:148-155), from barycentrics solved here;
* the grid of a cube mesh: dimensions and origin by hand; exact, positive
  values at voxels within a voxel of the box whose nearest cube feature is a
  vertex or an edge; a voxel below a face where the interior formula makes
  the nearest triangle a side face's and the value negative (derived by hand);
  the flood fill's positive values deep inside (the reference's stated
  assumption, ToolSDF.cu:510-514);
* the sampler: voxel-centre values, midpoints, clamping, the range test.
"""
import numpy as np
import pytest

import oracle


def exact_d2(p, a, b, c):
    """Exact squared distance point-triangle (float64, independent of the oracle)."""
    p, a, b, c = (np.asarray(x, dtype=np.float64) for x in (p, a, b, c))
    n = np.cross(b - a, c - a)
    n = n / np.linalg.norm(n)
    q = p - np.dot(p - a, n) * n
    # barycentric inside test of the projection
    T = np.array([b - a, c - a]).T
    lam, *_ = np.linalg.lstsq(T, q - a, rcond=None)
    if lam[0] >= 0 and lam[1] >= 0 and lam.sum() <= 1:
        return float(np.dot(p - q, p - q))

    def seg(p, u, v):
        t = np.clip(np.dot(p - u, v - u) / np.dot(v - u, v - u), 0, 1)
        d = p - (u + t * (v - u))
        return float(np.dot(d, d))
    return min(seg(p, a, b), seg(p, b, c), seg(p, c, a))


def reference_interior_d2(p, a, b, c):
    p, a, b, c = (np.asarray(x, dtype=np.float64) for x in (p, a, b, c))
    n = np.cross(b - a, c - a)
    n = n / np.linalg.norm(n)
    q = p - np.dot(p - a, n) * n
    T = np.array([b - a, c - a]).T
    (wb, wc), *_ = np.linalg.lstsq(T, q - a, rcond=None)
    wa = 1.0 - wb - wc
    r = a + wa * (b - a) + wb * (c - a)
    return float(np.dot(p - r, p - r)), (wa, wb, wc)


def test_point_triangle_regions():
    rng = np.random.default_rng(7)
    a, b, c = np.array([0.0, 0, 0]), np.array([1.0, 0.2, 0]), np.array([0.3, 0.9, 0.1])
    n_in = n_out = 0
    for _ in range(400):
        p = rng.uniform(-1.0, 2.0, 3)
        d2, _ = oracle.point_tri_d2(p, a, b, c)
        ref, (wa, wb, wc) = reference_interior_d2(p, a, b, c)
        if min(wa, wb, wc) > 1e-3:
            n_in += 1
            assert d2 == pytest.approx(ref, rel=2e-5, abs=1e-6)
        elif min(wa, wb, wc) < -1e-3:
            n_out += 1
            assert d2 == pytest.approx(exact_d2(p, a, b, c), rel=2e-5, abs=1e-6)
    assert n_in > 20 and n_out > 100
    # the interior formula is not the true projection: the centroid-free case
    d2, q = oracle.point_tri_d2([0.2, 0.2, 1.0], [0, 0, 0], [1, 0, 0], [0, 1, 0])
    assert d2 == pytest.approx(1.16, rel=1e-6) and q == pytest.approx([0.6, 0.2, 0.0], abs=1e-6)


def cube_mesh(lo, hi):
    """12 outward-facing triangles of the box [lo, hi]^3."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    V = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]])
    quads = [((0, 3, 2, 1), (0, 0, -1)), ((4, 5, 6, 7), (0, 0, 1)), ((0, 1, 5, 4), (0, -1, 0)),
             ((2, 3, 7, 6), (0, 1, 0)), ((1, 2, 6, 5), (1, 0, 0)), ((0, 4, 7, 3), (-1, 0, 0))]
    T, N = [], []
    for (i, j, k, l), n in quads:
        T += [(i, j, k), (i, k, l)]
        N += [n, n]
    return V, np.array(T, dtype=np.int32), np.array(N, dtype=np.float64)


def box_exact(p, lo, hi):
    """Signed distance of the box (negative inside)."""
    lo, hi = np.asarray(lo), np.asarray(hi)
    c, h = (lo + hi) / 2, (hi - lo) / 2
    q = np.abs(p - c) - h
    return np.linalg.norm(np.maximum(q, 0)) + min(q.max(), 0.0)


@pytest.fixture(scope="module")
def cube_sdf():
    lo, hi = (0.0, 0.0, 0.0), (4e-3, 3e-3, 2e-3)
    V, T, N = cube_mesh(lo, hi)
    return oracle.SdfGrid(V, T, N, padding=1e-3, voxels_per_axis=24), lo, hi


def test_cube_grid_layout(cube_sdf):
    s, lo, hi = cube_sdf
    f = np.float32
    pad = f(1e-3)
    blo = [f(x) - pad for x in lo]                 # float arithmetic, as ToolSDF::build
    bhi = [f(x) + pad for x in hi]
    ext = [f(h - l) for l, h in zip(blo, bhi)]
    vox = f(max(ext) / f(24))
    assert s.voxel == float(vox)
    assert s.origin == tuple(float(x) for x in blo)
    assert s.dims == tuple(min(256, int(f(e / vox)) + 2) for e in ext)


def test_cube_values_and_sign(cube_sdf):
    s, lo, hi = cube_sdf
    G = s.grid()
    nx, ny, nz = s.dims
    v = s.voxel
    o = np.array(s.origin)
    checked = 0
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = o + (np.array([i, j, k]) + 0.5) * v
                ex = box_exact(c, lo, hi)
                outside_axes = np.sum((c < np.array(lo)) | (c > np.array(hi)))
                # outside on two or more axes and within a voxel of the box:
                # in the scatter range of the triangles of the nearest edge or
                # vertex, whose closest points are exact; the sign from their
                # normals is outside
                if outside_axes >= 2 and ex < 0.95 * v:
                    checked += 1
                    assert G[k, j, i] == pytest.approx(ex, rel=1e-5, abs=1e-9), (i, j, k)
    assert checked > 20
    # half a voxel below the bottom face, away from the side edges the
    # interior formula of the bottom triangles overestimates the distance,
    # so the nearest triangle is a side face's (closest point on the edge
    # y = z = 0) and its normal -y makes the value negative:
    # -sqrt(dy^2 + dz^2), the reference's result for this voxel
    c = o + (np.array([6, 4, 3]) + 0.5) * v
    assert c[1] == pytest.approx(0.5 * v) and c[2] == pytest.approx(-0.5 * v)
    assert G[3, 4, 6] == pytest.approx(-np.hypot(c[1], c[2]), rel=1e-5)
    # deep inside the cube (beyond the scatter reach of every face) the flood
    # fill leaves positive values
    ci = ((np.array(lo) + np.array(hi)) / 2 - o) / v - 0.5
    i, j, k = (int(round(x)) for x in ci)
    assert box_exact(o + (np.array([i, j, k]) + 0.5) * v, lo, hi) < -3 * v
    assert G[k, j, i] > 0


def test_sampler(cube_sdf):
    s, lo, hi = cube_sdf
    G = s.grid()
    o, v = np.array(s.origin), s.voxel
    for (i, j, k) in [(3, 4, 5), (10, 2, 7), (0, 0, 0)]:
        c = o + (np.array([i, j, k]) + 0.5) * v
        assert s.sample(c) == pytest.approx(float(G[k, j, i]), rel=1e-12, abs=1e-15)
        m = c + np.array([0.5 * v, 0, 0])
        assert s.sample(m) == pytest.approx(0.5 * (float(G[k, j, i]) + float(G[k, j, i + 1])), rel=1e-12)
    # clamping: below the first voxel centre the edge value repeats
    c0 = o + 0.5 * v
    assert s.sample(c0 - 0.3 * v) == pytest.approx(float(G[0, 0, 0]), rel=1e-12)
    # range test on texel coordinates [0, n - 1]
    nx, ny, nz = s.dims
    assert s.in_range(o) and s.in_range(o + v * np.array([nx - 1, ny - 1, nz - 1]) - 1e-12)
    assert not s.in_range(o - np.array([1e-9, 0, 0]))
    assert not s.in_range(o + v * np.array([nx - 1, 0, 0]) + np.array([1e-9, 0, 0]))


def test_vertex_normals_and_wear_known_answers():
    """ContactSolver::initSDF's vertex normals: normalised face normals
    summed per triangle a vertex belongs to, then normalised.  A cube of two
    triangles per face: corner 0 lies in both triangles of its three faces,
    (-2, -2, -2) / |.|; corner 1 in one triangle of the bottom and front
    faces and both of the right face, (2, -1, -1) / sqrt(6).  A flat square
    whose face normals are (0, 0, 2) has every vertex normal (0, 0, 1); wear
    moves a vertex inward along it, vertices past the wear array stay."""
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    quads = [((0, 3, 2, 1), (0, 0, -1)), ((4, 5, 6, 7), (0, 0, 1)), ((0, 1, 5, 4), (0, -1, 0)),
             ((2, 3, 7, 6), (0, 1, 0)), ((1, 2, 6, 5), (1, 0, 0)), ((0, 4, 7, 3), (-1, 0, 0))]
    T, N = [], []
    for (i, j, k, l), n in quads:
        T += [(i, j, k), (i, k, l)]
        N += [n, n]
    vn = oracle.vertex_normals(8, T, np.array(N, float))
    assert np.allclose(vn[0], np.array([-1, -1, -1]) / np.sqrt(3), rtol=0, atol=1e-15)
    assert np.allclose(vn[1], np.array([2, -1, -1]) / np.sqrt(6), rtol=0, atol=1e-15)
    sq = np.array([[0, 0, 0], [2e-3, 0, 0], [2e-3, 2e-3, 0], [0, 2e-3, 0]], float)
    tris = [(0, 1, 2), (0, 2, 3)]
    nrm = np.array([[0, 0, 2.0], [0, 0, 2.0]])
    assert np.array_equal(oracle.vertex_normals(4, tris, nrm), np.tile([0.0, 0.0, 1.0], (4, 1)))
    worn = oracle.worn_vertices(sq, tris, nrm, [1e-4, 0.0, 2e-4])
    assert np.array_equal(worn, sq - np.array([[0, 0, 1e-4], [0, 0, 0], [0, 0, 2e-4], [0, 0, 0]]))
