"""The result formats' restatement (oracle/formats.py) against known answers.

The expected strings are written out by hand: base64 of FP32 1.0, 2.0, 3.0
(bytes 00 00 80 3f, 00 00 00 40, 00 00 40 40) with one, two and no padding
characters, the NDJSON message's sorted keys, and a two-particle VTK file
with a pressure of -(sxx + syy + szz) / 3 in MPa, a speed |v| and an
INACTIVE particle left out.
"""
import numpy as np

import formats


def test_base64_fp32_known_answers():
    assert formats.b64([1.0, 2.0, 3.0]) == "AACAPwAAAEAAAEBA"
    assert formats.b64([1.0]) == "AACAPw=="
    assert formats.b64([1.0, 2.0]) == "AACAPwAAAEA="
    assert formats.b64(np.zeros(0)) == ""


def test_ndjson_known_answer():
    line = formats.ndjson_particles(7, [[1, 2, 3], [9, 9, 9], [2, 3, 1]], [1.0, 5.0, 2.0], [0, 1, 2])
    assert line == (b'{"count":2,"positions_b64":"AACAPwAAAEAAAEBAAAAAQAAAQEAAAIA/","step":7,'
                    b'"temps_b64":"AACAPwAAAEA=","type":"particles"}\n')


def test_ndjson_nothing_when_all_inactive():
    assert formats.ndjson_particles(3, [[1, 2, 3]], [1.0], [1]) == b""


def test_vtk_known_answer():
    pos = np.array([[1e-3, 2e-3, 3e-3], [0, 0, 0], [4e-3, 5e-3, 6e-3]])
    vel = np.array([[3.0, 4.0, 0.0], [0, 0, 0], [-1.0, 0.0, 0.0]])
    stress = np.array([[-3e6, -3e6, -3e6, 0, 0, 0], [0] * 6, [1e6, 2e6, 3e6, 0, 0, 0]])
    text = formats.vtk_particles(pos, vel, [293.0, 300.0, 450.125], stress, [0, 1, 2])
    assert text == """# vtk DataFile Version 3.0
EdgePredict SPH Particles
ASCII
DATASET POLYDATA
POINTS 2 double
1.00000000e-03 2.00000000e-03 3.00000000e-03
4.00000000e-03 5.00000000e-03 6.00000000e-03

VERTICES 2 4
1 0
1 1

POINT_DATA 2
SCALARS Temperature_C double 1
LOOKUP_TABLE default
293.00
450.12

SCALARS Pressure_MPa double 1
LOOKUP_TABLE default
3.000000e+00
-2.000000e+00

SCALARS Velocity_m_s double 1
LOOKUP_TABLE default
5.000000e+00
1.000000e+00

SCALARS ParticleStatus int 1
LOOKUP_TABLE default
0
2

VECTORS Velocity double
3.000000e+00 4.000000e+00 0.000000e+00
-1.000000e+00 0.000000e+00 0.000000e+00
"""
