"""mpmw_export_ndjson_noalloc: the NDJSON result message built in the
library's static per-device scratch array instead of allocated buffers.

Checked through the C ABI against mpmw_export_ndjson (the same line, byte for
byte) on two calls in a row on different solvers, the second one smaller, so
the second call's message cannot borrow anything the first left in the
shared array; then a solver beyond the array's limit is refused with
MPMW_E_RANGE and no byte of the caller's buffer written.
"""
import ctypes as C

import numpy as np
import pytest

import mpmw
import workloads

pytestmark = pytest.mark.gpu


def _solver(cells, precision, seed, steps, inactive_every=0):
    cfg = workloads.scaled(workloads.drill_config1(), cells, steps=steps)
    P = dict(cfg["params"], precision=mpmw.FP64 if precision == "fp64" else mpmw.FP32)
    s = mpmw.Solver(**P)
    workloads.apply_to_solver(mpmw, s, dict(cfg, seed=seed))
    if inactive_every:
        d = s.get_particles()
        d["status"][::inactive_every] = mpmw.INACTIVE
        s.set_particles(d)
    s.run(steps, cfg["dt"])
    s.synchronize()
    return s


def _noalloc(s, step, cap=None):
    n = C.c_size_t(0)
    rc = mpmw.lib().mpmw_export_ndjson_noalloc(s.h, step, None, 0, C.byref(n))
    assert rc == 0, mpmw.lib().mpmw_last_error()
    buf = np.full((cap or n.value) + 64, 0xA5, dtype=np.uint8)
    rc = mpmw.lib().mpmw_export_ndjson_noalloc(s.h, step, buf.ctypes.data_as(C.c_char_p), buf.size,
                                               C.byref(n))
    assert rc == 0, mpmw.lib().mpmw_last_error()
    assert np.all(buf[n.value:] == 0xA5)          # nothing written past the line
    return buf[:n.value].tobytes()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_export_equals_export_two_calls(precision):
    big = _solver((96, 64, 40), precision, 0xED6E0101, 6, inactive_every=7)     # 1.97 M particles
    small = _solver((21, 13, 9), precision, 0xED6E0102, 4, inactive_every=5)    # 19,656 particles
    try:
        for s, step in ((big, 6), (small, 4)):
            want = s.export_ndjson(step)
            got = _noalloc(s, step)
            assert len(want) > 0 and got == want, (s.count(), len(got), len(want))
        # again in the other order: the smaller message first, then the larger
        for s, step in ((small, 4), (big, 6)):
            assert _noalloc(s, step) == s.export_ndjson(step)
    finally:
        big.close()
        small.close()


def test_noalloc_export_refuses_beyond_the_scratch_array():
    s = _solver((128, 64, 40), "fp32", 0xED6E0103, 1)                          # 2.62 M particles
    try:
        n = C.c_size_t(12345)
        buf = np.full(64, 0x5A, dtype=np.uint8)
        rc = mpmw.lib().mpmw_export_ndjson_noalloc(s.h, 1, buf.ctypes.data_as(C.c_char_p), buf.size,
                                                   C.byref(n))
        assert rc == mpmw.E_RANGE
        assert "scratch" in mpmw.lib().mpmw_last_error().decode()
        assert np.all(buf == 0x5A)
        # the allocating export still serves it
        assert s.export_ndjson(1).startswith(b'{"count":')
    finally:
        s.close()


def test_noalloc_export_all_inactive_writes_nothing():
    s = _solver((4, 4, 2), "fp32", 0xED6E0104, 1)
    try:
        d = s.get_particles()
        d["status"][:] = mpmw.INACTIVE
        s.set_particles(d)
        n = C.c_size_t(99)
        rc = mpmw.lib().mpmw_export_ndjson_noalloc(s.h, 0, None, 0, C.byref(n))
        assert rc == 0 and n.value == 0
    finally:
        s.close()
