#!/usr/bin/env python3
"""Benchmark of the MI355X MPM workpiece step (one GPU).

Workload (one "step" = one MPMWorkpieceSolver::step over the whole
workpiece): 256x128x64-cell Ti-6Al-4V slab, 8 particles per cell with seeded
jitter (16,777,216 particles), analytic two-flute helical drill (D=10 mm,
30 deg helix, 118 deg point) at 3000 rpm and 0.1 mm/rev, dt = 1e-8 s, FP32
bulk arithmetic.  Inputs are resident in HBM before the timed region.

Prints one JSON line: particle-steps/s, the dominant kernel's HBM roofline
fraction (HIP events on the solver's stream), and the CPU oracle on the
host cores as a reported baseline.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "edegpredict-engine-rotary_amd")
sys.path.insert(0, PKG)

import numpy as np  # noqa: E402

import mpmw  # noqa: E402
import serve  # noqa: E402
import workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle-step of the plain FP32 step (H_kin = 0, no
# external force), G2P of step n fused with P2G of step n+1:
#   particles: read 64 B (x 12, sigma 24, eps_p, D, sigma_res, T, status, m, V)
#              write 72 B (x 12, v 12, sigma 24, eps_p, rate, D, sigma_res, T, status)
#   grid: 96 B per occupied node (P2G flush 32, update read 32 + write 16, G2P read 16)
B_PARTICLE = 136.0
B_NODE_STEP = 96.0
# of which the fused G2P+P2G kernel moves: all particle bytes, the G2P node
# read (16 B) and the P2G flush (32 B)
B_NODE_G2P = 48.0
# the APIC transfer's fused pass also writes the affine matrix C (9 x 4 B)
B_PARTICLE_APIC = 136.0 + 36.0


class HipEvents:
    """hipEvent pairs on an arbitrary stream, through libamdhip64."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def create(self):
        e = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def occupied_nodes(cells):
    cx, cy, cz = cells
    return (cx + 1) * (cy + 1) * (cz + 1)


def host_cpu_info():
    """Cores this process may run on, the machine's count and its CPU model."""
    info = {"nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count(), "model": None,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return info


def cpu_baseline(cfg, budget_s, threads, runs=5):
    """The CPU FP64 oracle (restatement of the reference step) on the host
    cores, over a bounded sample of the same workload: one warm-up step,
    then `runs` timed runs of whole steps (each about budget_s / runs of
    work); the reported rate is the median run.  The serial leg does the
    same on one thread with `runs` single steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    import parity
    oracle.set_threads(threads)
    t0 = time.time()
    o, _ = parity.build_oracle(cfg)
    setup = time.time() - t0
    n = o.count()
    t0 = time.time()
    o.step(cfg["dt"])                      # warm-up step (includes the first grid rebuild)
    first = time.time() - t0
    per_run = max(1, min(20, int(budget_s / runs / max(first, 1e-3))))
    rates = []
    for _ in range(runs):
        t0 = time.time()
        for _ in range(per_run):
            o.step(cfg["dt"])
        rates.append(n * per_run / (time.time() - t0))
    oracle.set_threads(1)
    srates = []
    for _ in range(runs):
        t0 = time.time()
        o.step(cfg["dt"])
        srates.append(n / (time.time() - t0))
    oracle.set_threads(threads)
    info = host_cpu_info()
    return {
        "value": float(np.median(rates)), "unit": "particle-steps/s", "cores": threads, "kind": "port",
        "sample": f"{cfg['name']}: {n} particles, FP64 oracle, 1 warm-up step then the median of {runs} runs "
                  f"of {per_run} steps on {threads} OpenMP threads (setup {setup:.1f} s)",
        "runs": rates, "host": info,
        "serial": {"value": float(np.median(srates)), "unit": "particle-steps/s", "cores": 1,
                   "sample": f"median of {runs} single steps on one thread", "runs": srates},
    }


def run_mode(cfg, mode, precision, steps, warmup, breakdown_steps=25, graph=False):
    """Set up the workload in one solver and time `steps` steps after
    `warmup` untimed ones; HIP events on the solver's stream bracket the
    timed region, and library events bracket every launch of the fused
    particle pass inside it (the roofline's kernel duration).  Events around
    every kernel class would add two records per launch to the stream, which
    lengthens the step, so the other classes' averages come from
    `breakdown_steps` further steps after the timed region.  With `graph`
    the step's launches are replayed as HIP graphs and the timed region
    carries no per-kernel events (the kernel averages then come from the
    breakdown steps, stepped eagerly)."""
    P = dict(cfg["params"], precision=mpmw.FP64 if precision == "fp64" else mpmw.FP32,
             deterministic=1 if mode == "deterministic" else 0)
    s = mpmw.Solver(**P)
    s.set_graph(graph)
    t0 = time.time()
    workloads.apply_to_solver(mpmw, s, cfg)
    setup_s = time.time() - t0
    n = s.count()
    dt = cfg["dt"]
    for _ in range(warmup):
        s.step(dt)
    # untimed steps up to the next grid rebuild (every 25th step), so the
    # timed window starts on a rebuild step and holds ceil(steps / 25) of
    # them: never fewer than the workload's own share
    align = (-s.metrics().step) % 25
    for _ in range(align):
        s.step(dt)
    s.synchronize()
    first_timed = s.metrics().step
    ev = HipEvents()
    stream = s.stream()
    a, b = ev.create(), ev.create()
    s.set_kernel_timing(mpmw.TIMING_OFF if graph else mpmw.TIMING_PARTICLE_PASS)
    s.synchronize()
    t_wall = time.perf_counter()
    ev.record(a, stream)
    for _ in range(steps):
        s.step(dt)
    ev.record(b, stream)
    s.synchronize()
    wall = time.perf_counter() - t_wall
    ms_total = ev.elapsed_ms(a, b)
    kms, kn = s.kernel_times()
    fused_ms, fused_n = kms[mpmw.K_G2P_P2G], kn[mpmw.K_G2P_P2G]
    graph_stats = s.graph_stats()
    # per-class breakdown (not part of the timed region)
    s.set_kernel_timing(mpmw.TIMING_ALL)
    for _ in range(breakdown_steps):
        s.step(dt)
    s.synchronize()
    kms, kn = s.kernel_times()
    s.set_kernel_timing(mpmw.TIMING_OFF)
    # the result message of the final state (NDJSON particle line, built on
    # the device), timed as the C call into a preallocated host buffer
    nb = C.c_size_t(0)
    mpmw.lib().mpmw_export_ndjson(s.h, steps, None, 0, C.byref(nb))
    buf = np.empty(max(nb.value, 1), dtype=np.uint8)
    t_exp = []
    for _ in range(3):
        t = time.perf_counter()
        rc = mpmw.lib().mpmw_export_ndjson(s.h, steps, buf.ctypes.data_as(C.c_char_p), nb.value, C.byref(nb))
        t_exp.append(time.perf_counter() - t)
        assert rc == 0, mpmw.lib().mpmw_last_error()
    message = {"bytes": nb.value, "ms": min(t_exp) * 1e3,
               "GBs_to_host": nb.value / min(t_exp) / 1e9}
    diag = s.diagnostics()
    g = s.grid_info()
    s.close()
    names = ["g2p_p2g", "p2g", "grid_update", "diffuse", "sort", "reduce", "other"]
    return {
        "n": n, "ms_total": ms_total, "wall": wall, "setup_s": setup_s, "diag": diag,
        "window": {"alignment_steps": align, "first_timed_step": first_timed,
                   "rebuild_steps_timed": sum(1 for i in range(first_timed, first_timed + steps) if i % 25 == 0)},
        "grid_nodes": g.nx * g.ny * g.nz,
        "k_avg_ms": fused_ms / max(fused_n, 1), "k_launches": fused_n, "message": message,
        "graph": graph_stats if graph else None,
        "kernels": {names[i]: {"avg_ms": kms[i] / kn[i], "launches": kn[i]} for i in range(7) if kn[i]},
    }


def batched_bench(cfgs, precision, steps, graph=False, autotune=0, serial=None, cold=False):
    """The batched workload (64 workpieces of uneven size): wall time of the
    serial plain solves (one solver after another, each set up, stepped and
    synchronised) against the same set-ups stepped together in one batch.
    Set-up (generating each slab on the device) counts in both.  With
    `autotune` the batch tunes its launch configuration during its first
    steps (counted in the time); `serial` reuses an earlier serial time.
    With `cold` the batch runs twice and both times are reported: the first
    batch of the process also pays the device allocations of its 64
    members' state (the serial solves hold one member at a time), the
    second finds the allocator's pool warm, as a serving process does."""
    P = lambda c: dict(c["params"], precision=mpmw.FP64 if precision == "fp64" else mpmw.FP32)
    nsim = len(cfgs)
    parts = sum(int(np.prod(c["cells"])) * 8 for c in cfgs)
    if serial is None:
        parts = 0
        t0 = time.perf_counter()
        for c in cfgs:
            s = mpmw.Solver(**P(c))
            workloads.apply_to_solver(mpmw, s, c)
            parts += s.count()
            s.run(steps, c["dt"])
            s.synchronize()
            s.close()
        serial = time.perf_counter() - t0
    def run_batch():
        t0 = time.perf_counter()
        solvers = []
        b = mpmw.Batch()
        b.set_graph(graph)
        if autotune:
            b.set_autotune(autotune)
        for c in cfgs:
            s = mpmw.Solver(**P(c))
            workloads.apply_to_solver(mpmw, s, c)
            solvers.append(s)
            b.add(s)
        dts = np.array([c["dt"] for c in cfgs])
        for _ in range(steps):
            b.step(dts)
        b.synchronize()
        t = time.perf_counter() - t0
        graph_stats = b.graph_stats()
        tuning = b.tuning() if autotune else None
        b.close()
        for s in solvers:
            s.close()
        return t, graph_stats, tuning

    first = run_batch()[0] if cold else None
    batched, graph_stats, tuning = run_batch()
    return {
        "workload": f"{nsim} workpieces of the batched distribution, {steps} steps each",
        "particles": parts, "steps": steps,
        "serial_s": serial, "batched_s": batched,
        "serial_sims_per_hour": nsim / serial * 3600.0,
        "batched_sims_per_hour": nsim / batched * 3600.0,
        "speedup": serial / batched,
        "batched_particle_steps_per_s": parts * steps / batched,
        "first_batch": None if first is None else {
            "batched_s": first, "speedup": serial / first,
            "note": "the process's first batch, which also allocates its members' device state"},
        "graph": graph_stats if graph else None,
        "tuning": tuning,
    }


# F-tracking step (the reference's MPMSolver), FP32 bulk: bytes a particle
# must move per step.  P2G reads x (24), status + LOD (8), m, V (8), v (12),
# f_ext (12), C (36), T (4), F - I (36), Fp - I (36), backstress (24), eps_p,
# rate, damage, residual (16) = 216 and writes stress (24), eps_p, T, damage
# (12), Fp - I (36), status + LOD (8) = 80; G2P reads x (24), status + LOD
# (8), F - I (36), stress and damage for the chip pass (28) = 96 and writes x
# (24), v (12), C (36), F - I (36), rate (4), f_ext (12) = 124.  The grid
# (one node per 8 particles: 7 channels cleared, accumulated, read and 10
# written by the update, 3 read by G2P, 4 B each) adds ~12 B.
B_FTRACK = 216.0 + 80.0 + 96.0 + 124.0 + 12.0


def ftrack_bench(steps, warmup, dims=(512, 256, 128)):
    """The F-tracking step on a slab of particles at spacing dx / 2 (8 per
    cell of the reference's dx = 0.5 mm lattice), FP32 bulk with FP64
    islands, pressed by a rasterised cylindrical tool block (radius 5 mm,
    nodes every 0.5 mm) moving down at 5 m/s whose nodes sit in contact."""
    import ftrack
    sp = 2.5e-4
    mn = (0.0, 0.0, 0.0)
    mx = tuple((d - 1) * sp for d in dims)
    P = ftrack.default_params(jc_A=1098e6, jc_B=1092e6, jc_n=0.93, jc_C=0.014, jc_m=1.1, young=113.8e9,
                              poisson=0.342, density=4430.0, specific_heat=526.0, melt_temp=1878.0,
                              ambient_temp=293.0, fixture_thickness=0.0, lod_enabled=1,
                              lod_active_radius=0.006, lod_near_radius=0.025, max_particles=2 ** 31 - 1)
    s = ftrack.Solver(P)
    t0 = time.time()
    s.init_box(mn, mx, sp)
    setup = time.time() - t0
    n = s.count()
    cx, cy, top = 0.5 * mx[0], 0.5 * mx[1], mx[2]
    g = np.arange(-10, 11) * 5e-4
    pts = np.array([(cx + a, cy + b, top + 0.2e-3 + c) for a in g for b in g for c in np.arange(0, 20) * 5e-4
                    if a * a + b * b <= 25e-6])
    s.set_tool_nodes(pts, np.tile([0.0, 0.0, -5.0], (len(pts), 1)))
    s.set_tool_position(cx, cy, top)
    dt = 1e-8
    for _ in range(warmup):
        s.step(dt)
    s.synchronize()
    ev = HipEvents()
    a, b = ev.create(), ev.create()
    stream = s.stream()
    ev.record(a, stream)
    for _ in range(steps):
        s.step(dt)
    ev.record(b, stream)
    ms = ev.elapsed_ms(a, b)
    d = s.diagnostics()
    m = s.metrics()
    s.close()
    pps = n * steps / (ms / 1e3)
    return {"workload": f"slab {dims[0]}x{dims[1]}x{dims[2]} particles at 0.25 mm (8 per 0.5 mm cell), "
                        f"{len(pts)} tool nodes pressing at 5 m/s, FP32 with FP64 islands, dt 1e-8",
            "particles": n, "steps": steps, "warmup": warmup, "ms_per_step": ms / steps,
            "value": pps, "unit": "particle-steps/s", "setup_s": setup,
            "roofline": {"bound": "hbm", "B_alg_per_particle_step": B_FTRACK,
                         "achieved_GBs": pps * B_FTRACK / 1e9, "frac": pps * B_FTRACK / 1e9 / HBM_PEAK_GBS},
            "max_plastic_strain_proxy_max_stress": m["max_stress"], "diagnostics": d}


def _tet_tool_run(steps, warmup, precision, dims):
    """One coupled run (see tet_tool_bench); returns its timing, the contact
    results and the particle and tool-node fields the precision study
    compares."""
    import ftrack
    import toolfem
    sp = 2.5e-4
    mx = tuple((d - 1) * sp for d in dims)
    # temperatures in degC here, as the tool's (ambient 25, the 1878 K melt at 1604.85)
    P = ftrack.default_params(jc_A=1098e6, jc_B=1092e6, jc_n=0.93, jc_C=0.014, jc_m=1.1, young=113.8e9,
                              poisson=0.342, density=4430.0, specific_heat=526.0, melt_temp=1604.85,
                              ambient_temp=25.0, fixture_thickness=0.0, lod_enabled=1,
                              lod_active_radius=0.006, lod_near_radius=0.025, max_particles=2 ** 31 - 1,
                              precision=precision)
    s = ftrack.Solver(P)
    t0 = time.time()
    s.init_box((0.0, 0.0, 0.0), mx, sp)
    n = s.count()
    cx, cy, top = 0.5 * mx[0], 0.5 * mx[1], mx[2]
    mesh = toolfem.cylinder_tool_mesh(5e-3, 15e-3, 5e-4, base=(cx, cy, top - 1e-4), face_samples=5)
    tool = toolfem.Tool(toolfem.tool_params(young=600e9, poisson=0.22, spindle_damping=1.0))
    tool.set_mesh(mesh)
    tool.init_driven(0.2)
    c = toolfem.Contact(toolfem.contact_params(contact_radius=5e-4, jc_A=1098e6, jc_B=1092e6, jc_n=0.93,
                                               jc_C=0.014, jc_m=1.1, jc_T_ref=25.0, jc_T_melt=1604.85,
                                               work_melt_temp=1604.85, work_specific_heat=526.0,
                                               work_physical_density=4430.0))
    s.set_tool_position(cx, cy, top)
    setup = time.time() - t0
    dt, v = 1e-8, 5.0
    k = 0
    most = 0

    def one(timers=None):
        nonlocal k, most
        tool.set_spindle_state((0.0, 0.0, -v * dt * k), (0.0, 0.0, -v))
        t = time.perf_counter()
        tool.step(dt)
        toolfem.link_tool_nodes(s, tool)
        if timers is not None:
            tool.synchronize()
            t1 = time.perf_counter()
            timers["tool"] += t1 - t
            t = t1
        s.step(dt)
        if timers is not None:
            s.synchronize()
            t1 = time.perf_counter()
            timers["mpm"] += t1 - t
            t = t1
        c.resolve(s, tool, dt)
        most = max(most, c.results()["contacts"])
        if timers is not None:
            timers["contact"] += time.perf_counter() - t
        k += 1

    for _ in range(warmup):
        one()
    s.synchronize()
    tool.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    s.synchronize()
    tool.synchronize()
    wall = time.perf_counter() - t0
    timers = {"tool": 0.0, "mpm": 0.0, "contact": 0.0}
    nb = max(3, steps // 4)
    for _ in range(nb):
        one(timers)
    res = c.results()
    counts = tool.counts()
    parts = s.get_particles(fields=("pos", "temperature", "plastic_strain"))
    surf = tool.surface()
    out = {"particles": n, "steps": steps, "warmup": warmup, "ms_per_step": wall * 1e3 / steps,
           "value": n * steps / wall, "setup_s": setup,
           "breakdown_ms_per_step": {kk: vv * 1e3 / nb for kk, vv in timers.items()},
           "contact": dict(res, most_contacts=most), "counts": counts, "total_steps": k}
    c.close()
    tool.close()
    s.close()
    return out, parts, surf


def tet_tool_bench(steps, warmup, dims=(512, 256, 128), precision_study=True):
    """The coupled step with a deformable tool: the F-tracking slab of the
    ftrack leg, a corotational tet tool (a 10 mm diameter, 15 mm cylinder at
    0.5 mm: 49,680 tets, 9,703 tet nodes; 35,258 surface nodes, the
    boundary triangles' barycentric points embedded in the tets; E = 600
    GPa, nu = 0.22) driven down at 5 m/s through its spindle springs with
    its tip 0.1 mm into the slab, the tool's surface nodes rasterised into
    the MPM grid and in XPBD hash contact with the particles (cell = contact
    radius = 2h = 0.5 mm).  Per step, in the engine's order: tool step, tool
    nodes to the MPM solver, MPM step, contact (which reads back its totals,
    one synchronisation per step).  The precision study runs the same steps
    with FP64 particles and reports the FP32 run's per-field error against
    it (positions, plastic strain, temperature, contact force)."""
    import ftrack
    r32, p32, s32 = _tet_tool_run(steps, warmup, ftrack.FP32, dims)
    counts = r32.pop("counts")
    out = {"workload": f"F-tracking slab {dims[0]}x{dims[1]}x{dims[2]} particles at 0.25 mm + corotational tet "
                       f"tool ({counts[2]} tets, {counts[1]} tet nodes, {counts[0]} surface nodes, E 600 GPa, "
                       f"nu 0.22) pressed at 5 m/s + XPBD hash contact (cell 0.5 mm), FP32 particles, dt 1e-8",
           "unit": "particle-steps/s", **r32}
    if not precision_study:
        return out
    r64, p64, s64 = _tet_tool_run(steps, warmup, ftrack.FP64, dims)
    r64.pop("counts")

    def byid(p):
        o = np.argsort(p["id"], kind="stable")
        return {k: v[o] for k, v in p.items()}

    a, b = byid(p32), byid(p64)
    same = a["status"] == b["status"]

    def rel(x, y):
        d = np.abs(x - y).max() if x.size else 0.0
        return float(d / max(np.abs(y).max(), 1e-300))

    out["fp64"] = {"ms_per_step": r64["ms_per_step"], "value": r64["value"],
                   "breakdown_ms_per_step": r64["breakdown_ms_per_step"], "contact": r64["contact"]}
    out["precision_study"] = {
        "what": f"FP32 particles (FP64 positions and islands) against the FP64 run of the same "
                f"{r32['total_steps']} coupled steps, normwise max|d| / max|ref| over particles of equal status",
        "status_mismatch": int((~same).sum()),
        "positions": rel(a["pos"][same], b["pos"][same]),
        "plastic_strain": rel(a["plastic_strain"][same], b["plastic_strain"][same]),
        "temperature": rel(a["temperature"][same], b["temperature"][same]),
        "tool_node_force": rel(s32["force"], s64["force"]),
        "tool_node_temperature": rel(s32["temperature"], s64["temperature"]),
        "contact_force_total": abs(r32["contact"]["force"] - r64["contact"]["force"]) / max(abs(r64["contact"]["force"]), 1e-300),
    }
    return out


def measured_traffic():
    """HBM bytes per launch of the fused tile kernel on this workload, from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None if absent)."""
    path = os.path.join(ROOT, "profiles", "k_tile_drill_256x128x64_pmc.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def copy_calibration(mib=1024, reps=3):
    """A plain device copy of known size, the calibration kernel for the
    rocprofv3 FETCH_SIZE / WRITE_SIZE counters: `out = in * 1` over float32
    buffers (an elementwise kernel that reads and writes every byte once,
    which the runtime cannot turn into a DMA copy).  The ratio of the bytes a
    launch must move to what the counters report for it scales the counters
    of the solver's kernels to HBM bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"kernel": "elementwise mul (copy)", "read_bytes": 4 * n, "write_bytes": 4 * n,
                      "launches": reps, "GBs": 8 * n * reps / el / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=25)
    ap.add_argument("--modes", type=str, default="atomic,deterministic",
                    help="P2G modes to time; the first one gives `value`")
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--cells", type=str, default="256,128,64")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 timing of the same workload")
    ap.add_argument("--graph", type=int, default=1, help="1: also time the step replayed as HIP graphs")
    ap.add_argument("--autotune", type=int, default=2,
                    help="trials per candidate of the autotuned batched leg (0 skips it)")
    ap.add_argument("--tet-tool-steps", type=int, default=20,
                    help="steps of the coupled MPM + tet tool + hash contact leg (0 skips it)")
    ap.add_argument("--ftrack-steps", type=int, default=50,
                    help="steps of the F-tracking step's timing (0: skip it)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the plain copy kernel that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--batch-count", type=int, default=64, help="workpieces of the batched run (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--serve-seconds", type=float, default=4.0,
                    help="seconds of Poisson arrivals (20/s) for the served run (0: skip it)")
    ap.add_argument("--transfer", choices=["pic", "apic"], default="pic",
                    help="apic: the APIC transfer variant (atomic mode only)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    modes = [m.strip() for m in args.modes.split(",") if m.strip()]
    for m in modes:
        if m not in ("atomic", "deterministic"):
            ap.error(f"unknown mode {m}")

    cells = tuple(int(x) for x in args.cells.split(","))
    cfg = workloads.drill_config1()
    if cells != cfg["cells"]:
        cfg = workloads.scaled(cfg, cells)
    if args.transfer == "apic":
        # the deterministic mode is not available with the APIC transfer
        cfg = dict(cfg, transfer="apic")
        modes = [m for m in modes if m == "atomic"] or ["atomic"]
    runs = {m: run_mode(cfg, m, args.precision, args.steps, args.warmup) for m in modes}
    r = runs[modes[0]]
    n = r["n"]
    ms_step = r["ms_total"] / args.steps
    pps = n * args.steps / (r["ms_total"] / 1e3)
    nodes = occupied_nodes(cells)

    # dominant kernel: fused G2P + constitutive + chip + next-step P2G
    k_avg_ms = r["k_avg_ms"]
    b_part = B_PARTICLE_APIC if args.transfer == "apic" else B_PARTICLE
    bytes_launch = b_part * n + B_NODE_G2P * nodes
    achieved = bytes_launch / (k_avg_ms / 1e3) / 1e9
    b_alg = b_part + B_NODE_STEP * nodes / n

    out = {
        "metric": "MPM particle-steps/s at 1 GPU",
        "value": pps,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision + " (positions fp64)",
        "data": "synthetic",
        "config": {
            "workload": cfg["name"], "particles": n, "cells": list(cells),
            "grid_nodes": r["grid_nodes"], "occupied_nodes": nodes, "tool": "analytic two-flute drill",
            "dt": cfg["dt"], "mode": modes[0], "precision": args.precision, "transfer": args.transfer,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic() if args.transfer == "pic" else None,
            "kernel": ("k_tile<G2P,P2G> (fused G2P + J-C return + chip pass + next-step P2G)"
                       if args.transfer == "pic" else
                       "k_apic_tile<G2P,P2G> (APIC: 27-node gather, C update, J-C return, chip pass, next-step P2G)"),
            "algorithmic_bytes_per_launch": bytes_launch,
            "avg_launch_ms": k_avg_ms, "timed_launches": r["k_launches"],
            "step_level": {"B_alg_per_particle_step": b_alg,
                           "achieved_GBs": pps * b_alg / 1e9,
                           "frac": pps * b_alg / 1e9 / HBM_PEAK_GBS},
        },
        "kernels": r["kernels"],  # every class, from steps after the timed region
        "result_message": dict(r["message"], format="NDJSON particle line of the final state"),
        "binning": r["diag"],
        "timed_window": dict(r["window"], note="after the warm-up, untimed steps up to the next grid rebuild "
                                               "(every 25th step) so the timed steps include the rebuild "
                                               "steps at least in proportion"),
        "wall_s": r["wall"],
        "setup_s": r["setup_s"],
    }
    if args.graph:
        # the same atomic-mode workload with the step's launches replayed as
        # HIP graphs: the timed region has no kernel events, so the roofline
        # above keeps the eager run's fused-pass timing (the same kernel)
        q = run_mode(cfg, modes[0], args.precision, args.steps, args.warmup, graph=True)
        out["graph"] = {"value": n * args.steps / (q["ms_total"] / 1e3), "ms_per_step": q["ms_total"] / args.steps,
                        "mode": modes[0], "graph_stats": q["graph"], "timed_window": q["window"]}
    if args.precision == "fp32" and not args.no_fp64:
        # the same workload with every particle and grid field in FP64 (the
        # reference's own arithmetic), timed the same way
        q = run_mode(cfg, modes[0], "fp64", args.steps, args.warmup)
        out["fp64"] = {"value": n * args.steps / (q["ms_total"] / 1e3), "ms_per_step": q["ms_total"] / args.steps,
                       "mode": modes[0], "fused_kernel_avg_ms": q["k_avg_ms"], "kernels": q["kernels"],
                       "timed_window": q["window"]}
    for m in modes[1:]:
        q = runs[m]
        out[m] = {"value": n * args.steps / (q["ms_total"] / 1e3), "ms_per_step": q["ms_total"] / args.steps,
                  "fused_kernel_avg_ms": q["k_avg_ms"], "kernels": q["kernels"], "binning": q["diag"]}
    if args.batch_count > 0:
        bcfg = workloads.batch_config2(args.batch_count, steps=args.batch_steps)
        out["batched"] = batched_bench(bcfg, args.precision, args.batch_steps, cold=True)
        if args.graph:
            g = batched_bench(bcfg, args.precision, args.batch_steps, graph=True, serial=out["batched"]["serial_s"])
            out["batched"]["graph"] = {k: g[k] for k in ("serial_s", "batched_s", "speedup", "graph",
                                                         "batched_sims_per_hour")}
        if args.autotune > 0:
            # the autotuned batch (tuning inside its timed steps), against the
            # same serial time
            a = batched_bench(bcfg, args.precision, args.batch_steps, autotune=args.autotune,
                              serial=out["batched"]["serial_s"])
            out["batched"]["autotuned"] = {k: a[k] for k in ("batched_s", "speedup", "batched_sims_per_hour",
                                                             "tuning")}
    if args.ftrack_steps > 0:
        out["ftrack"] = ftrack_bench(args.ftrack_steps, 5)
    if args.tet_tool_steps > 0:
        out["tet_tool_contact"] = tet_tool_bench(args.tet_tool_steps, 3)
    if args.serve_seconds > 0:
        sv = {"workload": f"Poisson stream of 20 requests/s for {args.serve_seconds:g} s, sizes and parameters of "
                          "the batched distribution, 200-2000 steps each; a request's result is the NDJSON "
                          "particle message of its final state"}
        sv["served"] = serve.summary(serve.serve_batched(serve.request_stream(duration=args.serve_seconds),
                                                         args.precision))
        sv["serial"] = serve.summary(serve.serve_serial(serve.request_stream(duration=args.serve_seconds),
                                                        args.precision))
        sv["sims_per_hour_ratio"] = sv["served"]["sims_per_hour"] / sv["serial"]["sims_per_hour"]
        # At 20 requests/s the server keeps up, so its sims/hour is the arrival
        # rate; the same requests queued at once give each server's capacity.
        sat = {"workload": "the same requests, all queued at time 0"}
        sat["served"] = serve.summary(serve.serve_batched(
            serve.request_stream(duration=args.serve_seconds, backlog=True), args.precision))
        sat["serial"] = serve.summary(serve.serve_serial(
            serve.request_stream(duration=args.serve_seconds, backlog=True), args.precision))
        sat["sims_per_hour_ratio"] = sat["served"]["sims_per_hour"] / sat["serial"]["sims_per_hour"]
        sv["saturated"] = sat
        out["served"] = sv
    if not args.no_cpu:
        # the host cores this process may use (the OpenMP thread count when
        # the environment sets one: a share of a larger machine)
        omp = os.environ.get("OMP_NUM_THREADS", "")
        threads = int(omp) if omp.isdigit() and int(omp) > 0 else len(os.sched_getaffinity(0))
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_budget, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
