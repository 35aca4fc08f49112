"""CPU-side checks of the C ABI library: it loads, exports every symbol the
header declares, and its host-only functions agree with the oracle.  No
compute on a device happens here."""
import os
import re

import numpy as np
import pytest

import mpmw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "mpmw.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpmw_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = mpmw.lib()
    names = _declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(mpmw.EXPORTS)
    assert L.mpmw_abi_version() == 1


def test_ftrack_library_exports_every_declared_symbol():
    """include/mpms.h (the F-tracking solver): every declared entry point is
    exported and bound."""
    import ftrack
    src = open(os.path.join(ROOT, "include", "mpms.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(mpms_[a-z0-9_]+)\s*\(", src)))
    L = ftrack.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(ftrack.EXPORTS)


def test_ftrack_params_default_match_oracle_defaults():
    import ftrack
    import mpms as orc_ms
    p, o = ftrack.default_params(), orc_ms.default_params()
    for name, _ in orc_ms.Params._fields_:
        a, b = getattr(p, name), getattr(o, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name
    assert p.precision == ftrack.FP32 and p.island_det == 0.5 and p.newton_fp64 == 1


def test_ftrack_create_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    import ftrack
    with pytest.raises(mpmw.MPMError):
        ftrack.Solver()


def test_toolfem_library_exports_every_declared_symbol():
    """include/mpmc.h (the tool FEM and the hash contact): every declared
    entry point is exported and bound."""
    import toolfem
    src = open(os.path.join(ROOT, "include", "mpmc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(mpm[tc]_[a-z0-9_]+)\s*\(", src)))
    L = toolfem.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(toolfem.EXPORTS)


def test_toolfem_params_default_match_oracle_defaults():
    import mpmc as orc_mc
    import toolfem
    p, o = toolfem.tool_params(), orc_mc.tool_params()
    for name, _ in orc_mc.ToolParams._fields_:
        if name != "pad_":
            assert getattr(p, name) == getattr(o, name), name
    p, o = toolfem.contact_params(), orc_mc.contact_params()
    for name, _ in orc_mc.ContactParams._fields_:
        assert getattr(p, name) == getattr(o, name), name


def test_toolfem_create_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    import toolfem
    with pytest.raises(mpmw.MPMError):
        toolfem.Tool()
    with pytest.raises(mpmw.MPMError):
        toolfem.Contact()


def test_tool_mesh_generator_is_a_conforming_closed_tet_mesh():
    """cylinder_tool_mesh: positive tet volumes summing to the voxelised
    volume (cubes x h^3), every interior face shared by exactly two tets, the
    surface nodes exactly the nodes of the one-tet faces."""
    import toolfem
    m = toolfem.cylinder_tool_mesh(1.5e-3, 4e-3, 5e-4)
    X, el = m["tet_xyz"], m["elems"]
    D = np.stack([X[el[:, 1]] - X[el[:, 0]], X[el[:, 2]] - X[el[:, 0]], X[el[:, 3]] - X[el[:, 0]]], 2)
    vol = np.abs(np.linalg.det(D)) / 6
    assert vol.min() > 0 and len(el) % 6 == 0
    assert vol.sum() == pytest.approx(len(el) / 6 * 5e-4 ** 3, rel=1e-9)
    faces = np.sort(np.concatenate([el[:, [0, 1, 2]], el[:, [0, 1, 3]], el[:, [0, 2, 3]], el[:, [1, 2, 3]]]), 1)
    _, cnt = np.unique(faces, axis=0, return_counts=True)
    assert set(cnt.tolist()) <= {1, 2}
    assert len(m["surf_xyz"]) < len(X) and np.all(m["emb_w"][:, 0] == 1.0)


def test_params_default_match_oracle_defaults(orc):
    p = mpmw.default_params()
    o = orc.default_params()
    for name, _ in orc.Params._fields_:
        if name == "pad_":
            continue
        assert getattr(p, name) == getattr(o, name), name
    assert p.precision == mpmw.FP32 and p.deterministic == 0


def test_generator_matches_oracle_bit_exactly(orc):
    a = mpmw.gen_jittered_slab(0xED6E0001, 5, 4, 3, 2.5e-4, (1e-3, -2e-3, 0.5e-3), 4430.0, 7e-4)
    b = orc.gen_jittered_slab(0xED6E0001, 5, 4, 3, 2.5e-4, (1e-3, -2e-3, 0.5e-3), 4430.0, 7e-4)
    for k in ("pos", "mass", "volume", "id", "status"):
        assert np.array_equal(a[k], b[k]), k
    assert (a["status"] == mpmw.FIXED_BOUNDARY).any()


def test_generator_rejects_bad_sizes():
    with pytest.raises(mpmw.MPMError):
        mpmw.gen_jittered_slab(1, 0, 4, 3, 2.5e-4)


def test_create_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(mpmw.MPMError, match="HIP|device"):
        mpmw.Solver(**mpmw_params())


def mpmw_params():
    return {"cell_size": 2.5e-4}


def test_workload_tool_geometry():
    import workloads
    pts = workloads.rake_face_samples(0.0, 8e-3, 3.5e-3)
    assert pts.shape == (65 * 13, 3)
    # every sample lies on the rake plane through the edge, 5 degrees back
    n = np.array([np.cos(np.radians(5)), 0, np.sin(np.radians(5))])
    assert np.abs((pts - [0.0, 0.0, 3.5e-3]) @ n).max() < 1e-15
    c1 = workloads.drill_config1()
    assert np.prod(c1["cells"]) * 8 == 16_777_216


def test_backlog_stream_is_the_poisson_stream_queued_at_once():
    import serve
    live = serve.request_stream(duration=2.0)
    queued = serve.request_stream(duration=2.0, backlog=True)
    assert len(live) == len(queued) > 20
    assert all(0.0 < a.arrival < 2.0 for a in live) and all(q.arrival == 0.0 for q in queued)
    assert [a.steps for a in live] == [q.steps for q in queued]
    assert [a.cfg["cells"] for a in live] == [q.cfg["cells"] for q in queued]
