"""Voxel SDF tool (ToolSDF semantics) on the device against the oracle.

* The device build of the SDF (scatter of squared distances, band sign,
  host flood fill) equals the oracle's grid bit for bit, with the same
  dimensions, origin and voxel size, for the cutting-insert prism of the
  voxel workload, the reference application's parametric drill mesh, a
  closed sphere mesh of many small triangles (many
  triangles per voxel, ties between triangles) and the cube of the oracle's
  own SDF tests.
* A wedge workpiece cut by the insert mesh through its voxel SDF matches the
  oracle to the parity tolerances (FP64 1e-9, FP32 1e-5), grid included.
* Wear morphing rebuilds the SDF from the worn vertices bit for bit as the
  oracle does, also in the middle of a cut, which then stays within the
  parity tolerances.
* Argument errors are reported, not run.
"""
import numpy as np
import pytest

import mpmw
import oracle
import parity
import workloads

pytestmark = pytest.mark.gpu

GATED = ("pos", "temperature", "plastic_strain")


def sphere_mesh(r=1.2e-3, nu=24, nv=16, c=(0.3e-3, -0.2e-3, 0.1e-3)):
    """UV sphere with outward face normals."""
    c = np.asarray(c)
    V = [c + [0, 0, r], c - [0, 0, r]]
    for j in range(1, nv):
        th = np.pi * j / nv
        for i in range(nu):
            ph = 2 * np.pi * i / nu
            V.append(c + r * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)]))
    V = np.array(V)

    def ring(j, i):
        return 2 + (j - 1) * nu + (i % nu)
    T = []
    for i in range(nu):
        T.append((0, ring(1, i), ring(1, i + 1)))
        T.append((1, ring(nv - 1, i + 1), ring(nv - 1, i)))
    for j in range(1, nv - 1):
        for i in range(nu):
            a, b, d, e = ring(j, i), ring(j, i + 1), ring(j + 1, i), ring(j + 1, i + 1)
            T += [(a, d, e), (a, e, b)]
    T = np.array(T, dtype=np.int32)
    N = []
    for t in T:
        n = np.cross(V[t[1]] - V[t[0]], V[t[2]] - V[t[0]])
        n /= np.linalg.norm(n)
        if np.dot(n, V[t].mean(axis=0) - c) < 0:
            n = -n
        N.append(n)
    return V, T, np.array(N)


def cube_mesh(lo, hi):
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


def gpu_sdf(mesh, padding, vpa, precision=mpmw.FP32):
    s = mpmw.Solver(**dict(workloads.MATERIAL, fixture_thickness=0.0, precision=precision))
    d = mpmw.ToolDesc()
    for k, v in workloads.voxel_desc_values().items():
        setattr(d, k, v)
    s.set_tool_sdf_mesh(d, *mesh, padding, vpa)
    out = s.get_tool_sdf()
    s.close()
    return out


MESHES = {
    "insert": (lambda: workloads.prism_tool_mesh(17 * workloads.H + 2 * workloads.H), 0.002, 128),
    "sphere": (sphere_mesh, 0.4e-3, 96),
    "cube": (lambda: cube_mesh((0.0, 0.0, 0.0), (4e-3, 3e-3, 2e-3)), 1e-3, 24),
    # the reference application's parametric drill (main.cpp:83-233) at the
    # SDF build's own padding and resolution (ContactSolver.cu:1019)
    "parametric_drill": (lambda: mpmw.parametric_tool_mesh(mpmw.TOOLMESH_DRILL), 0.002, 128),
}


@pytest.mark.parametrize("name", sorted(MESHES))
def test_sdf_build_matches_oracle_bit_exact(name):
    make, pad, vpa = MESHES[name]
    mesh = make()
    ref = oracle.SdfGrid(*mesh, padding=pad, voxels_per_axis=vpa)
    G, origin, vox = gpu_sdf(mesh, pad, vpa)
    R = ref.grid()
    print(f"\n[sdf {name}] dims={ref.dims} voxel={ref.voxel:.4e} tris={len(mesh[1])} "
          f"negative={int((R < 0).sum())} differ={int((G.view(np.uint32) != R.view(np.uint32)).sum())}",
          flush=True)
    assert G.shape == R.shape == tuple(reversed(ref.dims))
    assert origin == ref.origin and vox == ref.voxel
    assert (R < 0).any()
    assert np.array_equal(G.view(np.uint32), R.view(np.uint32))


def test_worn_sdf_matches_oracle_bit_exact():
    """Wear morphing (ContactSolver::morphToolSDF): the insert mesh's
    vertices moved inward by per-vertex wear of up to 40 um (the last
    vertices given none, past the wear array), the SDF rebuilt on the device,
    equal bit for bit to the oracle's build of the oracle's worn vertices;
    a second morph starts again from the rest mesh."""
    V, T, N = workloads.prism_tool_mesh(17 * workloads.H + 2 * workloads.H)
    s = mpmw.Solver(**dict(workloads.MATERIAL, fixture_thickness=0.0))
    d = mpmw.ToolDesc()
    for k, v in workloads.voxel_desc_values().items():
        setattr(d, k, v)
    with pytest.raises(mpmw.MPMError):      # no mesh yet
        s.morph_tool_sdf(np.zeros(len(V)))
    s.set_tool_sdf_mesh(d, V, T, N, 0.002, 128)
    rest = s.get_tool_sdf()[0]
    rng = np.random.default_rng(7)
    for wear in (rng.uniform(0.0, 40e-6, len(V) - 2), rng.uniform(0.0, 10e-6, len(V))):
        s.morph_tool_sdf(wear)
        G, origin, vox = s.get_tool_sdf()
        ref = oracle.SdfGrid(oracle.worn_vertices(V, T, N, wear), T, N, padding=0.002, voxels_per_axis=128)
        R = ref.grid()
        print(f"\n[worn sdf] dims={ref.dims} max wear={wear.max():.2e} changed vs rest="
              f"{int((G.shape == rest.shape) and (G != rest).sum())}", flush=True)
        assert origin == ref.origin and vox == ref.voxel
        assert np.array_equal(G.view(np.uint32), R.view(np.uint32))
    s.close()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_wear_morph_mid_run_matches_oracle(prec, tol):
    """A voxel-tool cut whose tool wears after 15 steps: the solver morphs
    its SDF, the oracle is given the SDF of the same worn vertices, and both
    take 15 more steps."""
    cfg = voxel_case(30)
    o, _ = parity.build_oracle(cfg)
    s = parity.build_gpu(cfg, prec)
    for _ in range(15):
        o.step(cfg["dt"])
        s.step(cfg["dt"])
    V, T, N = cfg["tool"][3]
    wear = np.random.default_rng(3).uniform(0.0, 30e-6, len(V))
    s.morph_tool_sdf(wear)
    desc = oracle.ToolDesc()
    for k, v in cfg["tool"][1].items():
        setattr(desc, k, v)
    worn = oracle.SdfGrid(oracle.worn_vertices(V, T, N, wear), T, N, cfg["tool"][4], cfg["tool"][5])
    o.set_tool_sdf(worn, desc)
    G = s.get_tool_sdf()[0]
    assert np.array_equal(G.view(np.uint32), worn.grid().view(np.uint32))
    for _ in range(15):
        o.step(cfg["dt"])
        s.step(cfg["dt"])
    errs, mism = parity.compare_particles(s.get_particles(), o.get_particles())
    gm, om = s.metrics(), o.metrics()
    print(f"\n[worn tool {prec}] cf gpu/oracle {gm.contact_force:.6g}/{om.contact_force:.6g} "
          + " ".join(f"{k}={v:.2e}" for k, v in sorted(errs.items())), flush=True)
    assert mism == 0 and om.contact_force > 0.0
    for k in GATED:
        assert errs[k] <= tol, (k, errs[k])
    s.close()


def voxel_case(steps):
    cfg = sorted(workloads.batch_config2(), key=lambda c: np.prod(c["cells"]))[0]
    return workloads.voxel_tool_case(dict(cfg, steps=steps))


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_voxel_tool_matches_oracle(prec, tol):
    cfg = voxel_case(30)
    res = parity.run_case(cfg, prec)
    e = res["errors"]
    gm, om = res["metrics"]
    print(f"\n[voxel {cfg['name']} {prec}] cf gpu/oracle {gm.contact_force:.6g}/{om.contact_force:.6g} "
          + " ".join(f"{k}={v:.2e}" for k, v in sorted(e.items())), flush=True)
    assert om.contact_force > 0.0
    assert res["status_mismatch"] == 0 and res["grid_same"]
    for k in GATED + ("grid_mass", "grid_momentum"):
        assert e[k] <= tol, (k, e[k])
    assert gm.contact_force == pytest.approx(om.contact_force, rel=max(tol, 1e-6))


def test_voxel_tool_argument_errors():
    s = mpmw.Solver(**dict(workloads.MATERIAL, fixture_thickness=0.0))
    d = mpmw.ToolDesc()
    for k, v in workloads.voxel_desc_values().items():
        setattr(d, k, v)
    with pytest.raises(mpmw.MPMError):      # no SDF yet
        s.set_tool_analytic(d)
    with pytest.raises(mpmw.MPMError):
        s.get_tool_sdf()
    V, T, N = workloads.prism_tool_mesh(1e-3)
    with pytest.raises(mpmw.MPMError):      # empty mesh (ToolSDF::build refuses it)
        s.set_tool_sdf_mesh(d, V[:0], T[:0], N[:0], 0.002, 128)
    bad = T.copy()
    bad[0, 0] = len(V)
    with pytest.raises(mpmw.MPMError):      # vertex index out of range
        s.set_tool_sdf_mesh(d, V, bad, N, 0.002, 128)
    with pytest.raises(mpmw.MPMError):
        s.set_tool_sdf_mesh(d, V, T, N, 0.002, 0)
    s.set_tool_sdf_mesh(d, V, T, N, 0.002, 32)
    s.set_tool_analytic(d)                  # contact law update of the built SDF
    g, _, _ = s.get_tool_sdf()
    assert max(g.shape) <= 34 and (g < 0).any()
    s.close()
