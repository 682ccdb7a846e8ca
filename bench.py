#!/usr/bin/env python3
"""Benchmark: one RK3 step of the 9-species H2/O2/N2 WENO7 / 8th-order viscous /
finite-rate-chemistry path on a 256x128x128 curvilinear periodic block (FP64).

Prints ONE JSON line.  `value` is interior points x RK3 steps / wall time
(Mpoint-steps/s) with the state resident in HBM.  `roofline` covers the
kernel with the largest share of step time (HIP-event timing on the library's
stream); `cpu_baseline` times the oracle (the C restatement, OpenMP) on a
bounded sample of the same physics on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "corefl-cpc_amd"))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6

# Algorithmic FP64 words per interior point per call of each operator: each
# input plane read once and each output plane written once (nv = 14, ns = 9).
NV, NS = 14, 9
ALG_WORDS = {
    "zero_dq": NV,
    "viscous": (15 + 12) + (2 * NS + 13 + 3 * NS + 3) + (4 * NV - 3 + NV - 1),
    "chemistry": (2 * NS + 3) + (NS + 1),
    "inviscid_x": (NV + 6) + 2 * NV + 2 * NV,
    "inviscid_y": (NV + 6) + 2 * NV + 2 * NV,
    "inviscid_z": (NV + 6) + 2 * NV + 2 * NV,
    "update_rk": (3 * NV + 2) + (NV + 6 + NS),
    "properties": (4 + NS) + (5 + NS),
    "copy_qn": 2 * NV,
}
STEP_BYTES_PER_POINT = 8 * (3 * sum(ALG_WORDS[k] for k in ALG_WORDS if k != "copy_qn") + ALG_WORDS["copy_qn"])


PROFILE = REPO / "profiles" / "rhs_step_h2_9sp_256x128x128.md"
FP64_PROFILE = REPO / "profiles" / "fp64_rate_h2_9sp_256x128x128.md"
FP64_PEAK_TFLOPS = 78.6


def profile_flop(kernel: str):
    """FP64 FLOP per dispatch of a kernel from the committed counter summary
    (64 lanes x (2 FMA + MUL + ADD + TRANS) FP64 wave instructions)."""
    if not FP64_PROFILE.exists():
        return None
    for line in FP64_PROFILE.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 9 and cells[0] == kernel:
            try:
                return float(cells[5])
            except ValueError:
                return None
    return None
# kernels of each operator in the committed profile; the first one runs once per call
OP_KERNELS = {"viscous": ("k_viscous_flux", "k_zeta_grad", "k_periodic_flux", "k_viscous_derivative"),
              "inviscid_x": ("k_weno<9, true, 0",), "inviscid_y": ("k_weno<9, true, 1",),
              "inviscid_z": ("k_weno<9, true, 2",), "properties": ("k_properties",), "update_rk": ("k_update_rk",)}


def single_kernel(op: str, tiles) -> str | None:
    """The kernel an operator launches when it launches one (the WENO tile
    shape as the autotuner chose it: 0 = 64 / 16 faces per block, 1 = 2 x 32 for x, 8 x 32 for y and z)."""
    if op.startswith("inviscid_"):
        d = "xyz".index(op[-1])
        tx, ts = ((1, 64) if tiles[d] == 0 else (2, 32)) if d == 0 else (8, 16 << tiles[d])
        return f"k_weno<9, true, {d}, {tx}, {ts}>"
    return {"chemistry": "corefl_chem", "update_rk": "k_update_rk<9, false>", "properties": "k_properties<9>",
            "copy_qn": "k_copy", "zero_dq": "k_fill"}.get(op)


def profile_traffic(op: str, kernel: str | None = None):
    """HBM bytes per call of an operator (or of one named kernel) from the
    committed rocprofv3 summary (FETCH_SIZE and WRITE_SIZE per dispatch, scaled
    by the copy-kernel calibration there; kernels that run several times per
    call counted so)."""
    if not PROFILE.exists() or (kernel is None and op not in OP_KERNELS):
        return None
    rows = {}
    for line in PROFILE.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0].startswith("k_"):
            try:
                rows[cells[0]] = (int(cells[1]), float(cells[4]), float(cells[5]))
            except ValueError:
                continue
    def find(prefix):
        return next((v for k, v in rows.items() if k.startswith(prefix)), None)
    if kernel is not None:
        r = rows.get(kernel)
        return int((r[1] + r[2]) * 1024.0) if r else None
    first = find(OP_KERNELS[op][0])
    if first is None:
        return None
    total = 0.0
    for name in OP_KERNELS[op]:
        r = find(name)
        if r is None:
            return None
        total += (r[1] + r[2]) * 1024.0 * (r[0] / first[0])
    return int(total)


def cpu_cores() -> int:
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(sample=(256, 128, 64), steps=4):
    """Oracle (C restatement, OpenMP over points) on a bounded sample, in its
    -O3 -march=native build with FMA contraction (liboracle_fast.so, compiled
    on first use for the CPU it runs on)."""
    sys.path.insert(0, str(REPO / "oracle"))
    from corefl_amd.cases import reacting_box_case
    from oracle import Oracle, h2_tables

    case = reacting_box_case(*sample, seed=2024)
    o = Oracle(case, h2_tables(), fast=True)
    o.initialize()
    dt = 0.5 * o.cfl_dt()
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step(dt, True)
    el = time.perf_counter() - t0
    return {"value": case.points * steps / el / 1e6, "unit": "Mpoint-steps/s", "cores": cpu_cores(),
            "kind": "port",
            "sample": f"oracle C restatement (gcc -O3 -march=native, FMA contraction on), "
                      f"{sample[0]}x{sample[1]}x{sample[2]} block of the same physics, "
                      f"{steps} RK3 steps, {el:.1f} s, OpenMP threads={cpu_cores()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mx", type=int, default=256)
    ap.add_argument("--my", type=int, default=128)
    ap.add_argument("--mz", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--eager", action="store_true",
                    help="issue every operator from the host instead of replaying the recorded step graph")
    ap.add_argument("--workload", choices=("reacting", "multiblock", "serving"), default="reacting",
                    help="multiblock: twelve Air blocks (char. WENO5, inner faces), recorded step against "
                         "per-block launches; serving: a Poisson stream of 0-D H2/air reactor requests")
    ap.add_argument("--requests", type=int, default=10000, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=20000.0, help="serving: mean arrivals per second")
    ap.add_argument("--mb-width", type=int, default=0,
                    help="multiblock: nodes per side of every block (default: seeded sizes 32-192)")
    ap.add_argument("--tube-fraction", type=float, default=0.02, help="serving: share of shock-tube requests")
    ap.add_argument("--tube-t-end", type=float, default=2e-4, help="serving: shock-tube end time (s)")
    ap.add_argument("--tube-dt", type=float, default=1e-9, help="serving: shock-tube fixed step (s)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this implementation runs on one GPU; --gpus must be 1 (sharding blocks "
                                   "over several GPUs is not built)"}))
        sys.exit(2)

    import torch
    from corefl_amd import Solver, _abi, load_tables, reacting_box_case

    if args.workload == "multiblock":
        multiblock(args)
        return
    if args.workload == "serving":
        serving(args)
        return
    tables = load_tables()
    case = reacting_box_case(args.mx, args.my, args.mz, seed=2024, tables=tables)
    run = dict(case.run, step_graph=0 if args.eager else 1)
    solver = Solver(run, tables)
    solver.add_block(case)
    solver.initialize()
    dt = 0.5 * solver.cfl_dt()  # fixed dt: half the CFL-limited step of the initial field

    def sync():
        solver.sync()
        if torch.cuda.is_available():
            torch.cuda.synchronize()

    for _ in range(args.warmup):
        solver.step(dt)
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.step(dt)
    sync()
    elapsed = time.perf_counter() - t0
    pts = case.points
    value = pts * args.steps / elapsed / 1e6

    if args.profile_steps <= 0:
        print(json.dumps({"value": round(value, 3), "unit": "Mpoint-steps/s", "steps": args.steps}), flush=True)
        solver.close()
        return
    # live per-operator timing (HIP events on the library's stream): the same
    # step issued operator by operator (qn copy, three stages)
    solver.profile_begin()
    for _ in range(args.profile_steps):
        solver.op(_abi.OP_COPY_QN)
        for rk in range(3):
            solver.stage(rk, dt)
    prof = solver.profile_end()
    per_op = {}
    for name, (ms, calls) in prof.items():
        if calls == 0:
            continue
        avg_ms = ms / calls
        entry = {"avg_ms": round(avg_ms, 4), "calls": calls,
                 "share": round(ms / max(1e-30, sum(v[0] for v in prof.values())), 4)}
        if name in ALG_WORDS:
            b = ALG_WORDS[name] * 8 * pts
            entry["alg_bytes"] = b
            entry["achieved_GBs"] = round(b / (avg_ms * 1e-3) / 1e9, 1)
        per_op[name] = entry
    # the roofline of the dominant kernel: the single-kernel operator with the longest launch
    tiles = solver.tiles(0)
    single = {op: single_kernel(op, tiles) for op in per_op}
    dom = max((op for op, k in single.items() if k and op in ALG_WORDS), key=lambda op: per_op[op]["avg_ms"])
    d, kname = per_op[dom], single[dom]
    roofline = {"bound": "hbm", "kernel": kname, "operator": dom, "achieved": d["achieved_GBs"],
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(d["achieved_GBs"] / HBM_PEAK_GBS, 4),
                "traffic": profile_traffic(dom, kname), "alg_bytes_per_launch": d["alg_bytes"],
                "avg_launch_ms": d["avg_ms"],
                "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, copy-kernel scaled, "
                                  "profiles/rhs_step_h2_9sp_256x128x128.md"}
    flop = profile_flop(kname)
    if flop:
        tf = flop / (d["avg_ms"] * 1e-3) / 1e12
        roofline["fp64"] = {"achieved": round(tf, 2), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": round(tf / FP64_PEAK_TFLOPS, 4), "flop_per_launch": flop,
                            "flop_source": "rocprofv3 SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 per dispatch, "
                                           "profiles/fp64_rate_h2_9sp_256x128x128.md"}
    v = per_op.get("viscous")
    viscous = None
    if v:
        viscous = {"kernels": list(OP_KERNELS["viscous"]), "avg_ms": v["avg_ms"], "alg_bytes": v["alg_bytes"],
                   "achieved_GBs": v["achieved_GBs"], "frac": round(v["achieved_GBs"] / HBM_PEAK_GBS, 4),
                   "traffic": profile_traffic("viscous")}
    whole = STEP_BYTES_PER_POINT * value * 1e6 / 1e9
    out = {
        "metric": "Mpoint-steps/s FP64 9-sp WENO7 RK3 + % HBM roofline; p99 served-solve latency",
        "value": round(value, 3), "unit": "Mpoint-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3-D single-block curvilinear grid {args.mx}x{args.my}x{args.mz}, 9-species H2/O2/N2 "
                               "Li-2004 (19 reactions), characteristic WENO7 (inviscid_scheme 72), 8th-order "
                               "viscous, explicit chemistry, SSP-RK3, periodic faces",
                   "points": pts, "dt": dt, "step": "eager" if args.eager else "recorded step graph",
                   "weno_tiles": solver.tiles(0)},
        "roofline": roofline,
        "viscous_operator": viscous,
        "whole_step_hbm": {"alg_bytes_per_point_step": STEP_BYTES_PER_POINT, "achieved_GBs": round(whole, 1),
                           "frac": round(whole / HBM_PEAK_GBS, 4)},
        "per_op": per_op,
    }
    solver.close()
    if not args.no_cpu_baseline:
        cb = cpu_baseline()
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = round(value / cb["value"], 1)
    print(json.dumps(out), flush=True)


def multiblock(args):
    """Twelve Air blocks of seeded, uneven sizes (char. WENO5, 1-to-1 inner
    faces, fixed dt): the recorded step (one graph replay, blocks on
    concurrent streams) against the same step issued block by block."""
    import numpy as np
    import torch
    from corefl_amd import Solver, multiblock_case

    rng = np.random.default_rng(2025)
    widths = [sorted(int(v) for v in rng.integers(32, 193, size=3)), [int(v) for v in rng.integers(32, 97, size=2)],
              [int(v) for v in rng.integers(32, 97, size=2)]]
    if args.mb_width:  # every block the same size (the small end of the range: launch-bound blocks)
        widths = [[args.mb_width] * 3, [args.mb_width] * 2, [args.mb_width] * 2]
    blocks = multiblock_case(split=(3, 2, 2), widths=widths, seed=11, physics="air", rotate=False)
    pts = sum(c.points for c in blocks)
    res = {}
    dt = None
    for mode in ("eager", "graph"):
        run = dict(blocks[0].run, fixed_time_step=1, step_graph=1 if mode == "graph" else 0)
        s = Solver(run, None)
        for c in blocks:
            s.add_block(c)
        s.initialize()
        if dt is None:
            dt = 0.5 * s.cfl_dt()
        for _ in range(args.warmup):
            s.step(dt)
        s.sync()
        torch.cuda.synchronize() if torch.cuda.is_available() else None
        t0 = time.perf_counter()
        for _ in range(args.steps):
            s.step(dt)
        s.sync()
        res[mode] = time.perf_counter() - t0
        s.close()
    value = pts * args.steps / res["graph"] / 1e6
    print(json.dumps({
        "metric": "Mpoint-steps/s FP64 multiblock (recorded step vs per-block launches)",
        "value": round(value, 3), "unit": "Mpoint-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(res["graph"] / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "12 Air blocks (perfect gas), characteristic WENO5 + SSP-RK3, 1-to-1 inner faces, "
                               "outflow outer faces, fixed dt", "block_widths": widths, "points": pts},
        "eager": {"value": round(pts * args.steps / res["eager"] / 1e6, 3),
                  "ms_per_step": round(res["eager"] / args.steps * 1e3, 3)},
        "speedup_over_eager": round(res["eager"] / res["graph"], 3)}), flush=True)


TUBE_SEQ_SCALE = 20


def serving(args):
    """A stream of requests arriving as a seeded Poisson process, each either a
    0-D constant-volume H2/air reactor (T0 900-1600 K, p0 1-10 atm, equivalence
    ratio 0.5-2, to 2 ms, adaptive ROK4E with RTol 1e-4 / ATol 1e-8) or, with
    probability --tube-fraction, a 1-D reactive shock tube (64-4096 nodes over
    0.1 m, the same mixture at T0 / p0 in the driven three quarters, N2 at T0
    and 10 p0 in the driver, closed ends, WENO7 characteristic, Wu's operator
    splitting with explicit chemistry sub-steps at the fixed --tube-dt, to
    --tube-t-end: the reference's reactiveShockTube scheme).  Reactors and
    tubes go to their own servers, each driven by its own host thread on its
    own stream, so their launches overlap on the GPU.  Latency = completion
    minus arrival on the host clock.  The baselines: the same servers with one
    request in flight at a time (sequential per-request launches), timed on a
    sample of each kind, and the oracle's ROK4E reactor (C restatement,
    -O3 -march=native build) on the host cores, timed on a sample of the
    reactor requests."""
    import threading

    import numpy as np
    from corefl_amd import ReactorServer, TubeServer, h2_air_mass_fractions, load_tables

    tables = load_tables()
    mw = np.array(tables.mw[:9])
    rng = np.random.default_rng(2026)
    n = args.requests
    T0 = rng.uniform(900.0, 1600.0, n)
    p0 = rng.uniform(1.0, 10.0, n) * 101325.0
    phi = rng.uniform(0.5, 2.0, n)
    is_tube = rng.uniform(size=n) < args.tube_fraction
    cells = rng.integers(64, 4097, n)
    arrive = np.cumsum(rng.exponential(1.0 / args.rate, n))
    req = dict(t_end=2e-3, dt=1e-8, scheme="rok4e", adaptive=True, rtol=1e-4, atol=1e-8)
    n2 = np.zeros(9)
    n2[8] = 1.0

    def tube_args(i):
        y = h2_air_mass_fractions(phi[i], mw)
        return dict(n_cells=int(cells[i]), length=0.1, t_end=args.tube_t_end, n_left=int(cells[i]) // 4,
                    left=(T0[i], 10.0 * p0[i], 0.0, n2), right=(T0[i], p0[i], 0.0, y), bc=("symmetry", "symmetry"))

    rsrv = ReactorServer(tables, budget_bytes=256 << 20, steps_per_launch=32)
    tsrv = TubeServer(tables, run=dict(inviscid_scheme=72, temporal_scheme=4, fixed_time_step=1, dt=args.tube_dt),
                      budget_bytes=2 << 30, steps_per_progress=64)
    # warm-up (module load, the chemistry kernel compiled for the mechanism, first launches)
    w = rsrv.submit(1200.0, 101325.0, h2_air_mass_fractions(1.0, mw), **req)
    rsrv.run_until_done([w])
    w = tsrv.submit(**{**tube_args(0), "n_cells": 64, "n_left": 16, "t_end": 64 * args.tube_dt})
    tsrv.run_until_done([w])
    done = np.full(n, np.nan)
    steps = np.zeros(n, dtype=np.int64)
    t0 = time.perf_counter()

    def drive(kind):
        idx = np.flatnonzero(is_tube == (kind == "tube"))
        ticket_of, nxt, finished = {}, 0, 0
        shown = time.perf_counter()
        while finished < idx.size:
            if time.perf_counter() - shown > 30.0:  # progress of long runs, on stderr
                shown = time.perf_counter()
                print(f"serving {kind}s: {finished} of {idx.size} done at {shown - t0:.0f} s", file=sys.stderr,
                      flush=True)
            now = time.perf_counter() - t0
            while nxt < idx.size and arrive[idx[nxt]] <= now:
                i = idx[nxt]
                if kind == "tube":
                    t = tsrv.submit(**tube_args(i))
                else:
                    t = rsrv.submit(T0[i], p0[i], h2_air_mass_fractions(phi[i], mw), **req)
                ticket_of[t] = i
                nxt += 1
            if kind == "tube":
                pending = tsrv.progress()
                stamp = time.perf_counter() - t0
                for t in tsrv.ready():
                    r = tsrv.poll(t)
                    i = ticket_of.pop(t)
                    done[i], steps[i] = stamp, r["steps"]
                    finished += 1
            else:
                pending = rsrv.progress()
                stamp = time.perf_counter() - t0
                for t, r in rsrv.drain(max(n, 1)).items():
                    i = ticket_of.pop(t)
                    done[i], steps[i] = stamp, r["steps"]
                    finished += 1
            if pending == 0 and nxt < idx.size:
                time.sleep(max(0.0, arrive[idx[nxt]] - (time.perf_counter() - t0)))

    threads = [threading.Thread(target=drive, args=(k,)) for k in ("reactor", "tube")]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    wall = time.perf_counter() - t0
    lat = done - arrive
    # sequential baselines: one request in flight at a time, on a sample of each kind
    # a tube's sequential cost is timed to 1/TUBE_SEQ_SCALE of its end time and scaled (fixed dt: the
    # work grows with the step count), so the baseline stays within minutes
    seq = {}
    for kind, k_max in (("reactor", 20), ("tube", 2)):
        idx = np.flatnonzero(is_tube == (kind == "tube"))[:k_max]
        ts = time.perf_counter()
        for i in idx:
            if kind == "tube":
                t = tsrv.submit(**{**tube_args(i), "t_end": args.tube_t_end / TUBE_SEQ_SCALE})
                tsrv.run_until_done([t])
            else:
                t = rsrv.submit(T0[i], p0[i], h2_air_mass_fractions(phi[i], mw), **req)
                rsrv.run_until_done([t])
        per = (time.perf_counter() - ts) / max(idx.size, 1)
        seq[kind] = (per * (TUBE_SEQ_SCALE if kind == "tube" else 1), int(idx.size))
    rsrv.close()
    tsrv.close()
    n_t = int(is_tube.sum())
    n_r = n - n_t
    solves = n / wall
    seq_wall = n_r * seq["reactor"][0] + n_t * seq["tube"][0]

    def pct(mask):
        if not mask.any():
            return None
        v = lat[mask]
        return {"p50": round(float(np.percentile(v, 50)) * 1e3, 2), "p99": round(float(np.percentile(v, 99)) * 1e3, 2),
                "max": round(float(v.max()) * 1e3, 2)}

    tube_pts = float((cells[is_tube] * steps[is_tube]).sum())
    cpu = serving_cpu_anchor(T0[~is_tube][:48], p0[~is_tube][:48], phi[~is_tube][:48], mw, req)
    print(json.dumps({
        "metric": "served reactor and shock-tube solves/s and p50/p99 latency (FP64)",
        "value": round(solves, 1), "unit": "solves/s", "n_gpus": 1, "steps": None, "warmup": 1,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n} requests with Poisson arrivals at {args.rate:g}/s: {n_r} 0-D constant-volume "
                               "H2/air reactors (Li-2004; T0 900-1600 K, p0 1-10 atm, phi 0.5-2, to 2 ms, adaptive "
                               f"ROK4E RTol 1e-4 / ATol 1e-8) and {n_t} 1-D reactive shock tubes (64-4096 nodes over "
                               "0.1 m, N2 driver at 10 p0, closed ends, WENO7 characteristic, Wu operator splitting "
                               f"with explicit chemistry sub-steps, fixed dt {args.tube_dt:g} s, to "
                               f"{args.tube_t_end:g} s)",
                   "requests": n, "reactors": n_r, "tubes": n_t},
        "latency_ms": pct(np.ones(n, bool)), "latency_ms_reactors": pct(~is_tube), "latency_ms_tubes": pct(is_tube),
        "wall_s": round(wall, 3),
        "tube_mpoint_steps_per_s": round(tube_pts / wall / 1e6, 2) if n_t else None,
        "mean_steps": {"reactor": round(float(steps[~is_tube].mean()), 1) if n_r else None,
                       "tube": round(float(steps[is_tube].mean()), 1) if n_t else None},
        "max_steps": {"reactor": int(steps[~is_tube].max()) if n_r else None,
                      "tube": int(steps[is_tube].max()) if n_t else None},
        "sequential": {k: {"s_per_request": round(v[0], 4), "sample": v[1],
                           **({"timed_to": f"t_end / {TUBE_SEQ_SCALE}, scaled"} if k == "tube" else {})}
                       for k, v in seq.items()},
        "speedup_over_sequential": round(seq_wall / wall, 1),
        "cpu_baseline": cpu}), flush=True)


def serving_cpu_anchor(T0, p0, phi, mw, req):
    """The oracle's adaptive ROK4E reactor (the reference's const_volume_reactor
    scheme, C restatement in its -O3 -march=native build) on a sample of the
    reactor requests: seconds per request on one core, and the solves/s of all
    host cores working through the sample (one request per thread; ctypes
    releases the interpreter lock during the call)."""
    import ctypes as C
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    sys.path.insert(0, str(REPO / "oracle"))
    from corefl_amd import h2_air_mass_fractions
    from oracle import PD, h2_tables, load

    tb, lib = h2_tables(), load(fast=True)
    r_u = 8314.462618

    def one(i):
        y0 = h2_air_mass_fractions(phi[i], mw)
        rho = p0[i] / (r_u * (y0 / mw).sum() * T0[i])
        T = C.c_double(T0[i])
        y = np.ascontiguousarray(np.concatenate([y0, np.zeros(11)]))
        rej = C.c_long(0)
        t = time.perf_counter()
        lib.oc_cv_reactor_rok4e(C.byref(tb), rho, C.byref(T), y.ctypes.data_as(PD), req["dt"], req["t_end"], 1,
                                req["rtol"], req["atol"], C.byref(rej))
        return time.perf_counter() - t

    n = len(T0)
    single = [one(i) for i in range(min(n, 8))]
    cores = cpu_cores()
    t = time.perf_counter()
    with ThreadPoolExecutor(cores) as ex:
        list(ex.map(one, range(n)))
    wall = time.perf_counter() - t
    return {"value": round(n / wall, 1), "unit": "solves/s", "cores": cores, "kind": "port",
            "s_per_request_one_core": round(float(np.mean(single)), 4),
            "sample": f"{n} of the reactor requests (adaptive ROK4E RTol {req['rtol']:g} / ATol {req['atol']:g} to "
                      f"{req['t_end']:g} s) through the oracle's C restatement (gcc -O3 -march=native), one "
                      f"request per thread on {cores} threads, {wall:.1f} s; one-core latency from the first "
                      f"{len(single)}"}



if __name__ == "__main__":
    main()
