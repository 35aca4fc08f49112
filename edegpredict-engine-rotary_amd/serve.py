"""Serving loop: simulation requests admitted into a continuously stepped batch.

Requests arrive as a Poisson stream (the served workload: sizes and
parameters of the batched distribution, step counts uniform in [200, 2000]).
The server admits arrived requests, first come first served, while the
device-state budget allows; every admitted request becomes a solver with its
state resident on the device and joins one mpmw batch that the loop steps
continuously, so requests that arrive while others run share the device with
them (continuous batching, not a queue).  A request leaves the batch when it
has taken its steps; its solver's memory goes back to the library's pools.
The reference solver steps one workpiece per call; this loop has no
counterpart there.

The serial baseline serves the same stream with one plain solver at a time.
A request's result is the NDJSON particle message of its final state
(StdoutIPCExporter::exportParticles' format, built on the device); its
time-to-result ends when that message is on the host.

    python serve.py --duration 4 --rate 20
prints one JSON line with p50/p99 time-to-result and sims/hour for both.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import mpmw  # noqa: E402
import workloads  # noqa: E402


class Request:
    def __init__(self, rid, arrival, cfg):
        self.rid = rid
        self.arrival = arrival          # seconds after the stream starts
        self.cfg = cfg
        self.steps = cfg["steps"]
        self.solver = None
        self.done_steps = 0
        self.admitted = None
        self.finished = None
        self.metrics = None
        self.particles = None
        self.payload = None             # the result message (NDJSON particle line)

    @property
    def n(self):
        cx, cy, cz = self.cfg["cells"]
        return cx * cy * cz * 8


def request_stream(seed=workloads.SEED_SERVE, rate=20.0, duration=4.0, step_lo=200, step_hi=2000,
                   backlog=False):
    """Poisson arrivals at `rate` per second over `duration` seconds.  With
    `backlog` the same requests are all queued at time 0: a server's
    makespan then measures its capacity rather than the arrival rate."""
    rng = np.random.default_rng(seed)
    out, t, i = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        steps = int(rng.integers(step_lo, step_hi + 1))
        out.append(Request(i, 0.0 if backlog else t, workloads.sample_workpiece(rng, i, seed, steps)))
        i += 1
    return out


def state_bytes(req, precision):
    """Upper bound of the device bytes a request holds while resident (the
    admission budget's unit).  Counts every particle field the request's set-up
    makes live (positions FP64; backstress with kinematic hardening, the
    affine matrix with the APIC transfer) in both buffer sets, the binning
    arrays and the sort's temporary storage, the grid (8 accumulator channels,
    node record, mass, T*m) and the tile windows at 5/4 of the starting box
    (a grid grows by 5/4 when the workpiece outgrows it), the voxel SDF of a
    mesh tool, and the pool's size classes (at most 1/8 over a request)."""
    r = 8 if precision == "fp64" else 4
    cfg = req.cfg
    P = cfg["params"]
    apic = cfg.get("transfer") in ("apic", "apic_textbook")
    n = req.n
    reals = 17 + (6 if P.get("kinematic_hardening", 0.0) > 0.0 else 0) + (9 if apic else 0)
    per_particle = 2 * (3 * 8 + reals * r + 2 * 4) + 5 * 4 + 16
    cx, cy, cz = cfg["cells"]
    nodes = (cx + 10) * (cy + 10) * (cz + 10) * 5 // 4
    tiles = (cx // 8 + 3) * (cy // 8 + 3) * (cz // 4 + 3) * 5 // 4
    window = 847 if apic else 405
    total = n * per_particle + nodes * 14 * r + tiles * 8 * window * r
    if cfg["tool"][0] == "voxel":
        vox = int(cfg["tool"][5])
        total += 2 * 4 * vox ** 3
    return int(total * 9 // 8) + (4 << 20)


def _new_solver(req, precision, deterministic):
    P = dict(req.cfg["params"], precision=mpmw.FP64 if precision == "fp64" else mpmw.FP32,
             deterministic=1 if deterministic else 0)
    s = mpmw.Solver(**P)
    workloads.apply_to_solver(mpmw, s, req.cfg)
    return s


def _finish(req, t0, keep, payload):
    """Complete a request: its result message (the NDJSON particle line of
    the final state, built on the device) and metrics; the time-to-result
    ends when the message is on the host."""
    if payload == "ndjson":
        req.payload = req.solver.export_ndjson(req.steps)
    req.finished = time.perf_counter() - t0
    req.metrics = req.solver.metrics()
    if keep:
        req.particles = req.solver.get_particles()
    req.solver.close()
    req.solver = None


def serve_batched(requests, precision="fp32", deterministic=False, budget=64 << 30, max_slots=256,
                  keep_results=False, payload="ndjson"):
    """Serve the stream with continuous batching; returns the requests with
    admission and completion times (seconds after the start)."""
    pending = sorted(requests, key=lambda q: q.arrival)
    active = []
    used = 0
    b = mpmw.Batch()
    t0 = time.perf_counter()
    nxt = 0
    while nxt < len(pending) or active:
        now = time.perf_counter() - t0
        # admit arrived requests in order while the budget and slots allow
        while nxt < len(pending) and pending[nxt].arrival <= now and len(active) < max_slots:
            q = pending[nxt]
            need = state_bytes(q, precision)
            if active and used + need > budget:
                break
            q.solver = _new_solver(q, precision, deterministic)
            q.admitted = time.perf_counter() - t0
            b.add(q.solver)
            active.append(q)
            used += need
            nxt += 1
        if not active:
            wait = pending[nxt].arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            continue
        b.step([q.cfg["dt"] for q in active])
        for q in active:
            q.done_steps += 1
        done = [q for q in active if q.done_steps >= q.steps]
        if done:
            b.synchronize()
            for q in done:
                b.remove(q.solver)
                _finish(q, t0, keep_results, payload)
                used -= state_bytes(q, precision)
            active = [q for q in active if q.solver is not None]
    b.close()
    return requests


def serve_serial(requests, precision="fp32", deterministic=False, keep_results=False, payload="ndjson"):
    """Serve the stream one plain solver at a time, first come first served."""
    t0 = time.perf_counter()
    for q in sorted(requests, key=lambda r: r.arrival):
        wait = q.arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        q.solver = _new_solver(q, precision, deterministic)
        q.admitted = time.perf_counter() - t0
        q.solver.run(q.steps, q.cfg["dt"])
        q.solver.synchronize()
        q.done_steps = q.steps
        _finish(q, t0, keep_results, payload)
    return requests


def summary(requests):
    ttr = np.array([q.finished - q.arrival for q in requests])
    makespan = max(q.finished for q in requests)
    work = sum(q.n * q.steps for q in requests)
    return {
        "requests": len(requests),
        "p50_time_to_result_s": float(np.percentile(ttr, 50)),
        "p99_time_to_result_s": float(np.percentile(ttr, 99)),
        "makespan_s": makespan,
        "sims_per_hour": len(requests) / makespan * 3600.0,
        "particle_steps_per_s": work / makespan,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--duration", type=float, default=4.0, help="seconds of arrivals")
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--deterministic", action="store_true")
    ap.add_argument("--seed", type=int, default=workloads.SEED_SERVE)
    ap.add_argument("--no-serial", action="store_true")
    ap.add_argument("--backlog", action="store_true", help="queue every request at time 0")
    ap.add_argument("--payload", choices=["ndjson", "none"], default="ndjson",
                    help="result message of each request (NDJSON particle line)")
    args = ap.parse_args()
    out = {"workload": f"Poisson stream, {args.rate:g} requests/s for {args.duration:g} s, sizes of the batched "
                       f"distribution, 200-2000 steps each" + (", all queued at time 0" if args.backlog else ""),
           "precision": args.precision}
    reqs = request_stream(args.seed, args.rate, args.duration, backlog=args.backlog)
    out["served"] = summary(serve_batched(reqs, args.precision, args.deterministic, payload=args.payload))
    if not args.no_serial:
        reqs = request_stream(args.seed, args.rate, args.duration, backlog=args.backlog)
        out["serial"] = summary(serve_serial(reqs, args.precision, args.deterministic, payload=args.payload))
        out["sims_per_hour_ratio"] = out["served"]["sims_per_hour"] / out["serial"]["sims_per_hour"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
