"""CPU restatement of the reference's result formats (test infrastructure only).

Only tests/ import this module; the product library never does.

* ndjson_particles: StdoutIPCExporter::exportParticles
  (src/StdoutIPCExporter.cpp:70-113).  The particles whose status is not
  INACTIVE, in index order; positions cast to FP32 (x, y, z interleaved) and
  temperatures cast to FP32, each array's little-endian bytes in base64
  (include/Base64.h:13-52 is the standard RFC 4648 alphabet with '='
  padding, which Python's base64 module implements), in one JSON object
  dumped by nlohmann::json (object keys in sorted order, no spaces), then a
  newline.  Nothing when no particle is exported (:79).
* vtk_particles: VTKExporter::exportParticles(filename, particles)
  (src/VTKExporter.cpp:252-336), with the pressure of getParticlesForExport
  (src/MPMWorkpieceSolver.cu:870-894), -(sxx + syy + szz) / 3.  iostream's
  std::scientific with precision p prints as printf's %.pe, std::fixed as
  %.pf.
"""
import base64
import json

import numpy as np

INACTIVE = 1


def b64(a):
    return base64.b64encode(np.ascontiguousarray(a).astype("<f4").tobytes()).decode("ascii")


def ndjson_particles(step, pos, temperature, status):
    live = np.asarray(status) != INACTIVE
    count = int(live.sum())
    if count == 0:
        return b""
    msg = {"type": "particles", "step": int(step), "count": count,
           "positions_b64": b64(np.asarray(pos, dtype=np.float64)[live].reshape(-1)),
           "temps_b64": b64(np.asarray(temperature, dtype=np.float64)[live])}
    return (json.dumps(msg, sort_keys=True, separators=(",", ":")) + "\n").encode("ascii")


def vtk_particles(pos, vel, temperature, stress, status):
    status = np.asarray(status)
    live = np.nonzero(status != INACTIVE)[0]
    m = len(live)
    out = ["# vtk DataFile Version 3.0\n", "EdgePredict SPH Particles\n", "ASCII\n", "DATASET POLYDATA\n",
           f"POINTS {m} double\n"]
    out += ["%.8e %.8e %.8e\n" % tuple(pos[i]) for i in live]
    out.append(f"\nVERTICES {m} {2 * m}\n")
    out += [f"1 {k}\n" for k in range(m)]
    out.append(f"\nPOINT_DATA {m}\n")
    out.append("SCALARS Temperature_C double 1\nLOOKUP_TABLE default\n")
    out += ["%.2f\n" % temperature[i] for i in live]
    out.append("\nSCALARS Pressure_MPa double 1\nLOOKUP_TABLE default\n")
    for i in live:
        s = stress[i]
        p = -((s[0] + s[1]) + s[2]) / 3.0
        out.append("%.6e\n" % (p / 1e6))
    out.append("\nSCALARS Velocity_m_s double 1\nLOOKUP_TABLE default\n")
    for i in live:
        vx, vy, vz = (float(c) for c in vel[i])
        out.append("%.6e\n" % np.sqrt((vx * vx + vy * vy) + vz * vz))
    out.append("\nSCALARS ParticleStatus int 1\nLOOKUP_TABLE default\n")
    out += ["%d\n" % status[i] for i in live]
    out.append("\nVECTORS Velocity double\n")
    out += ["%.6e %.6e %.6e\n" % tuple(vel[i]) for i in live]
    return "".join(out)
