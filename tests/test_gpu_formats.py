"""Result formats from the device against their restatement (oracle/formats.py).

* NDJSON particle message, built on the device (compaction + FP32 cast +
  base64 of positions and temperatures): byte for byte equal to the
  restatement of the same downloaded state, for 0 to 5 exported particles
  (every base64 padding case of the temperatures), INACTIVE particles mixed
  in, a stepped rake case in both precisions, and a 1.05 M-particle slab
  after steps (there also through the size-independent round trip: decoding
  the strings gives back the FP32 positions and temperatures).
* Legacy VTK file: text equal to the restatement.
* Buffer contract: a NULL buffer reports the length, a short one fails.
"""
import base64
import json

import numpy as np
import pytest

import formats
import mpmw
import parity
import workloads

pytestmark = pytest.mark.gpu


def expected_ndjson(s, step):
    p = s.get_particles()
    return formats.ndjson_particles(step, p["pos"], p["temperature"], p["status"])


def small_solver(statuses, seed=1):
    rng = np.random.default_rng(seed)
    n = len(statuses)
    cfg = parity.small_case(steps=0)
    s = mpmw.Solver(**dict(cfg["params"], precision=mpmw.FP32))
    d = mpmw.make_arrays(n)
    d["pos"][:] = rng.uniform(0.0, 2e-3, (n, 3))
    d["vel"][:] = rng.normal(0.0, 2.0, (n, 3))
    d["temperature"][:] = rng.uniform(293.0, 900.0, n)
    d["stress"][:] = rng.normal(0.0, 3e8, (n, 6))
    d["mass"][:] = 1e-9
    d["volume"][:] = 1e-13
    d["density"][:] = 4430.0
    d["id"][:] = np.arange(n)
    d["status"][:] = statuses
    s.set_particles(d)
    return s


@pytest.mark.parametrize("statuses", [[1, 1], [0], [2, 1, 0], [0, 0, 1, 3], [4, 0, 1, 0, 2], [0, 0, 0, 0, 0]])
def test_ndjson_small_matches_restatement(statuses):
    s = small_solver(statuses)
    got = s.export_ndjson(12)
    assert got == expected_ndjson(s, 12)
    if all(st == 1 for st in statuses):
        assert got == b""
    s.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_ndjson_after_steps(prec):
    cfg = parity.small_case(steps=30)
    s = parity.build_gpu(cfg, prec)
    for _ in range(30):
        s.step(cfg["dt"])
    assert s.export_ndjson(30) == expected_ndjson(s, 30)
    s.close()


def test_ndjson_large_slab_round_trip():
    cfg = workloads.scaled(workloads.drill_config1(), (128, 64, 16), steps=3)
    s = parity.build_gpu(cfg, "fp32")
    for _ in range(3):
        s.step(cfg["dt"])
    line = s.export_ndjson(3)
    p = s.get_particles()
    live = p["status"] != mpmw.INACTIVE
    msg = json.loads(line)
    assert line.endswith(b"\n") and msg["count"] == int(live.sum()) > 1_000_000
    pos = np.frombuffer(base64.b64decode(msg["positions_b64"]), dtype="<f4").reshape(-1, 3)
    tmp = np.frombuffer(base64.b64decode(msg["temps_b64"]), dtype="<f4")
    assert np.array_equal(pos, p["pos"][live].astype(np.float32))
    assert np.array_equal(tmp, p["temperature"][live].astype(np.float32))
    assert line == formats.ndjson_particles(3, p["pos"], p["temperature"], p["status"])
    s.close()


def test_ndjson_buffer_contract():
    s = small_solver([0, 0, 0])
    n = mpmw.C.c_size_t(0)
    assert mpmw.lib().mpmw_export_ndjson(s.h, 5, None, 0, mpmw.C.byref(n)) == 0
    want = expected_ndjson(s, 5)
    assert n.value == len(want)
    short = mpmw.C.create_string_buffer(n.value - 1)
    assert mpmw.lib().mpmw_export_ndjson(s.h, 5, short, n.value - 1, mpmw.C.byref(n)) == mpmw.E_RANGE
    s.close()


def test_vtk_matches_restatement(tmp_path):
    s = small_solver([0, 1, 2, 3, 4, 0, 1])
    path = tmp_path / "particles_000012.vtk"
    s.export_vtk(str(path))
    p = s.get_particles()
    want = formats.vtk_particles(p["pos"], p["vel"], p["temperature"], p["stress"], p["status"])
    assert path.read_text() == want
    s.close()
