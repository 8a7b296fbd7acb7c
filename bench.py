#!/usr/bin/env python3
"""Benchmark of the MI355X CA time step on the 800^3 directional-solidification workload.

A "step" is one CA time step (Nucleation + CellCapture, one cycle of the reference's loop,
src/runCA.cpp:290-342) over the whole 804 x 804 x 802-cell domain. The run first advances untimed
to --start-cycle (6000: the middle of the run, where the solidification front is fully developed;
the run ends at cycle 11000), then takes --warmup untimed steps, then times exactly --steps steps
from the device-resident state (termination check every 1000 steps as in the reference); value =
active-cell updates per second, i.e. the sum over the timed steps of the frontier size (cells past
their liquidus time, the reference's numSteer) divided by the device time of those steps (HIP
events on the solver's stream).

roofline: the dominant work is CellCapture (k_claim, k_win, k_apply, k_settle and k_rebuild, the
reference's CellCapture kernel). Its algorithmic bytes are 137 B per active-cell update + 125 B per
capture + 1 B per solidification (the reference layout's bytes: own state, the 26 critical lengths
and the new cell's octahedron), divided by the kernels' event-timed durations over the same cycles
(a second solve of the deck reaches the same state and steps the window again with an event pair
around every kernel). traffic is the measured HBM bytes per step of the same kernels over the same
cycles (rocprofv3 FETCH_SIZE and WRITE_SIZE, calibrated on a plain copy) from the committed summary
profiles/capture_kernels_800_hbm_traffic.md; null when that summary covers other cycles.
With EXACA_PROFILE_REGIONS set, the two windows are bracketed by roctxProfilerResume/Pause for
rocprofv3 --selected-regions.

cpu_baseline: the oracle (the C restatement of the reference's algorithm) on the same window: it
starts from the GPU's state at the step before the window (cell types, GrainIDs, half-diagonals,
double octahedron centres; the oracle derives the undercooling and critical lengths as the
reference holds them) and steps the same cycles, serially and with OpenMP on every core this
process may use, as the reference's Serial and OpenMP backends run CellCapture; value is the
OpenMP rate, the serial one is reported beside it, and the serial run's end state must equal the
GPU's (serial_state_equals_gpu).

ensemble: --ensemble process-map requests (256: the batched-ensemble configuration's instance
count; directional solidification, cube sides log-uniform in
[32, 256] cells, G, R and nucleation density log-uniform, one seed each, Poisson arrivals at
50 req/s) solved three ways on the same GPU: all at once in the batched ensemble, served as they
arrive (continuous batching under a 200 GiB budget), and one after another on the plain path.
solves/s, latency percentiles and whether every result is identical across the three runs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

METRIC = "active-cell updates/s + HBM GB/s at 800³; ensemble solves/s, p99 vs plain loop"
HBM_PEAK_GBPS = 8000.0
TRAFFIC_PROFILE = REPO / "profiles" / "capture_kernels_800_hbm_traffic.md"


def measured_traffic() -> tuple[int | None, tuple[int, int] | None]:
    """HBM bytes per step of the CellCapture kernels from the committed rocprofv3 summary, and the
    cycles (first, last) its dispatches cover."""
    if not TRAFFIC_PROFILE.exists():
        return None, None
    traffic = window = None
    for ln in TRAFFIC_PROFILE.read_text().splitlines():
        if ln.startswith("traffic_bytes_per_step:"):
            traffic = int(ln.split(":")[1])
        if ln.startswith("traffic_cycles:"):
            a, b = ln.split(":")[1].split("-")
            window = (int(a), int(b))
    return traffic, window


def fast_leg(deck: Path, w0: int, steps: int) -> dict:
    """The fast build (floating-point contraction allowed; judged by grain statistics, not bit
    parity) on the same window: device time per step and its active-cell update rate."""
    import exaca_amd as ea

    with ea.Solver(deck, fast=True) as sol:
        if w0 > 0:
            sol.run(max_cycles=w0)
        q0 = sol.query()
        cycle, ms = sol.run(max_cycles=w0 + steps)
        q1 = sol.query()
    n = max(cycle - q0["cycle"], 1)
    return {"ms_per_step": ms / n, "value": (q1["frontier"] - q0["frontier"]) / (ms / 1e3),
            "captures_per_step": (q1["captures"] - q0["captures"]) / n,
            "build": "libexaca_amd_fast.so: -ffp-contract=fast (fp32 divide and sqrt still correctly rounded)"}


def host_cpu() -> tuple[str, int]:
    """CPU model and the cores this process may use."""
    model = "unknown"
    try:
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def cpu_baseline(deck: Path, state: dict, cycle: int, steps: int, after: dict) -> dict:
    """The oracle (the C restatement of the reference's algorithm) on the GPU line's own window:
    it starts from the GPU's state at `cycle` (the step before the timed window) and steps the same
    `steps` cycles, serially (the reference's Serial backend) and with OpenMP on every core this
    process may use (the reference's OpenMP backend: atomic frontier appends and compare-exchange
    captures). The serial run must end in the GPU's state (`after`); its CellCapture time is the
    reference log's bucket (src/runCA.cpp:294-311)."""
    sys.path.insert(0, str(REPO / "oracle"))
    import numpy as np

    import exaca_amd as ea
    import oracle as orc

    model, ncpu = host_cpu()
    threads = min(ncpu, int(os.environ.get("OMP_NUM_THREADS", ncpu)))
    legs = {}
    parity = None
    for name, nthreads in (("serial", 1), ("openmp", threads)):
        t0 = time.perf_counter()
        sim = orc.OracleSim.from_state(deck, ea.DATA_DIR, cycle, state["ct_raw"], state["gid"], state["dl"],
                                       state["centre"], state["nloc"], state["ntime"])
        sim.set_threads(nthreads)
        load_s = time.perf_counter() - t0
        end, cap_s = sim.run(max_cycles=cycle + steps)
        d = sim.dims()
        legs[name] = {"updates_per_s": d["frontier"] / cap_s, "capture_s": cap_s, "threads": nthreads,
                      "steps": end - cycle, "frontier": d["frontier"], "captures": d["captures"],
                      "import_s": load_s}
        if name == "serial":
            parity = bool(np.array_equal(sim.field(orc.CELLTYPE), after["ct"])
                          and np.array_equal(sim.field(orc.GRAINID), after["gid"]))
        print(f"[bench] cpu {name} ({nthreads} threads): {end - cycle} steps, CellCapture {cap_s:.2f} s, "
              f"{d['frontier']} active-cell updates", file=sys.stderr, flush=True)
        sim.close()
    omp = legs["openmp"]
    return {
        "value": omp["updates_per_s"],
        "unit": "active-cell updates/s",
        "cores": omp["threads"],
        "kind": "port",
        "sample": f"the GPU line's own window: cycles {cycle + 1}-{cycle + steps} of {deck.name} (804x804x802), "
                  f"started from the GPU's state at cycle {cycle}; the oracle's CellCapture run as the reference's "
                  f"OpenMP backend on {omp['threads']} threads ({omp['capture_s']:.2f} s for {omp['frontier']} "
                  f"active-cell updates) and serially ({legs['serial']['capture_s']:.2f} s)",
        "serial": legs["serial"],
        "openmp": omp,
        "cpu_model": model,
        "nproc": ncpu,
        "serial_state_equals_gpu": parity,
    }


def ensemble_leg(n: int) -> dict:
    from exaca_amd import decks, serve

    with tempfile.TemporaryDirectory() as d:
        reqs = decks.process_map_requests(n, d, rate=50.0)
        cells = sum((r.nx + 2) * (r.ny + 2) * (r.nz + 2) for r in reqs)
        batched = serve.run_served(reqs, realtime=False)
        print(f"[bench] batched ensemble: {n} solves in {batched['makespan_s']:.2f} s", file=sys.stderr, flush=True)
        served = serve.run_served(reqs, realtime=True)
        print(f"[bench] served: p50 {served['p50_s']:.3f} s p99 {served['p99_s']:.3f} s", file=sys.stderr, flush=True)
        seq = serve.run_sequential(reqs)
        print(f"[bench] sequential plain path: {seq['wall_s']:.2f} s", file=sys.stderr, flush=True)
    identical = all(seq["digests"][k] == batched["digests"][k] == served["digests"][k] for k in seq["digests"])
    return {
        "requests": n,
        "cells": cells,
        "batched_solves_per_s": batched["solves_per_s"],
        "served_solves_per_s": served["solves_per_s"],
        "served_p50_s": served["p50_s"],
        "served_p99_s": served["p99_s"],
        "sequential_solves_per_s": seq["solves_per_s"],
        "sequential_p50_s": seq["p50_s"],
        "sequential_p99_s": seq["p99_s"],
        "speedup_served_vs_sequential": served["solves_per_s"] / seq["solves_per_s"],
        "speedup_batched_vs_sequential": batched["solves_per_s"] / seq["solves_per_s"],
        # the sequential loop's host initialisation runs one request at a time (threads inside a
        # request only for large domains); without it the loop would take wall_s - init_s
        "sequential_init_s": seq["init_s"],
        "sequential_gpu_s": seq["gpu_s"],
        "speedup_batched_vs_sequential_gpu_only": batched["solves_per_s"] * seq["gpu_s"] / n,
        "blocks_per_slot": batched["blocks_per_slot"],
        "blocks_per_slot_served": served["blocks_per_slot"],
        "peak_resident": batched["peak_resident"],
        "peak_resident_served": served["peak_resident"],
        "failed": sorted(set(batched["failed"]) | set(served["failed"])),
        "identical_outputs": identical,
        "arrivals": "Poisson 50 req/s (served, sequential); all at t = 0 (batched)",
    }


def roctx():
    """roctxProfilerResume/Pause when EXACA_PROFILE_REGIONS is set (for rocprofv3 --selected-regions:
    only the timed windows are profiled); no-ops otherwise."""
    if not os.environ.get("EXACA_PROFILE_REGIONS"):
        return (lambda: None), (lambda: None)
    import ctypes

    lib = ctypes.CDLL("librocprofiler-sdk-roctx.so")
    return (lambda: lib.roctxProfilerResume(0)), (lambda: lib.roctxProfilerPause(0))


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--start-cycle", type=int, default=6000,
                    help="untimed advance before the warm-up, into the run's steady front")
    ap.add_argument("--deck", default=str(REPO / "examples" / "Inp_DirSolidification_800.txt"))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fast", action="store_true", help="skip the fast build's timing of the same window")
    ap.add_argument("--ensemble", type=int, default=256, help="process-map requests (0: skip)")
    ap.add_argument("--copy-probe-gib", type=int, default=4, help="size of the HBM copy calibration (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs one GPU: the whole CA domain lives on a single MI355X (asked for {args.gpus})",
              file=sys.stderr)
        return 2

    import exaca_amd as ea

    resume, pause = roctx()
    deck = Path(args.deck)
    t0 = time.perf_counter()
    sol = ea.Solver(deck)
    init_s = time.perf_counter() - t0
    print(f"[bench] initialised {deck.name}: {sol.X}x{sol.Y}x{sol.nz} cells in {init_s:.1f} s", file=sys.stderr,
          flush=True)
    copy_gbps = None
    if args.copy_probe_gib > 0:
        copy_bytes = args.copy_probe_gib << 30
        copy_ms = ea.copy_probe(copy_bytes, 10)
        copy_gbps = 2 * copy_bytes / (copy_ms / 1e3) / 1e9

    # untimed advance into the steady front, then the warm-up, then exactly --steps timed steps
    w0 = args.start_cycle + args.warmup
    if w0 > 0:
        sol.run(max_cycles=w0)
    q0 = sol.query()
    if q0["cycle"] != w0:
        raise RuntimeError(f"the run ended at cycle {q0['cycle']} before the timed window (cycle {w0})")
    t1 = time.perf_counter()
    resume()
    cycle, gpu_ms = sol.run(max_cycles=w0 + args.steps)
    pause()
    wall = time.perf_counter() - t1
    q1 = sol.query()
    steps = cycle - q0["cycle"]
    window = f"cycles {q0['cycle'] + 1}-{cycle} after an untimed advance to cycle {args.start_cycle} " \
             f"and {args.warmup} warm-up steps"
    print(f"[bench] timed {steps} steps ({window}) in {gpu_ms:.3f} ms device / {wall:.3f} s wall; "
          f"{q1['frontier'] - q0['frontier']} active-cell updates, {q1['captures'] - q0['captures']} captures",
          file=sys.stderr, flush=True)
    frontier = q1["frontier"] - q0["frontier"]
    value = frontier / (gpu_ms / 1e3)
    ncell, nuclei = sol.ncell, q1["nuclei"]
    sol.close()

    # The roofline kernels, event-timed one by one over the same window: a second solve of the same
    # deck reaches the same state (the run is deterministic) and steps the same cycles with an event
    # pair around every kernel. Its counters must equal the timed run's. The CPU baseline starts
    # from its state before the window and must end in its state after it.
    sol = ea.Solver(deck)
    if w0 > 0:
        sol.run(max_cycles=w0)
    state = sol.front_state() if not args.no_cpu else None
    resume()
    prof = sol.profile_kernels(steps)
    pause()
    after = {"ct": sol.field(ea.FIELD_CELLTYPE), "gid": sol.field(ea.FIELD_GRAINID)} if not args.no_cpu else None
    sol.close()
    for k in ("frontier", "captures", "solidified"):
        if prof[k] != q1[k] - q0[k]:
            raise RuntimeError(f"event-timed replay of the window differs in {k}: {prof[k]} vs {q1[k] - q0[k]}")
    kms = prof["ms"]
    nsteps = max(prof["steps"], 1)
    # CellCapture = claim + winners + apply + settle + the rebuild of the next list (the reference's pass 1
    # and pass 2)
    cap_ms = sum(v for k, v in kms.items() if k != "k_nucleate")
    bytes_alg = 137 * prof["frontier"] + 125 * prof["captures"] + 1 * prof["solidified"]
    achieved = bytes_alg / (cap_ms / 1e3) / 1e9 if cap_ms > 0 else 0.0
    traffic, traffic_window = measured_traffic()
    if traffic_window != (q0["cycle"] + 1, cycle):
        traffic = None  # the committed counter profile covers other cycles: not comparable
    step_us = 1e3 * cap_ms / nsteps
    measured = traffic / (step_us * 1e3) if traffic and step_us > 0 else None

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "active-cell updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": gpu_ms / max(steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32 (reference types per quantity)",
        "data": "synthetic: the deck's own seeded substrate, nuclei and orientations",
        "config": {
            "workload": "directional solidification 800^3 (G = R = 1e6, N = 10, substrate 1/25 um^2, "
                        f"nuclei 1e13 m^-3), {window}",
            "deck": deck.name,
            "cells": ncell,
            "nuclei": nuclei,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "CellCapture = k_claim + k_win + k_apply + k_settle + k_rebuild (one step)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic,
            "traffic_source": str(TRAFFIC_PROFILE.relative_to(REPO)) if traffic else None,
            "avg_launch_us": step_us,
            "profiled_steps": prof["steps"],
            "profiled_window": f"cycles {q0['cycle'] + 1}-{cycle}",
            "bytes_per_update": 137,
            "bytes_per_capture": 125,
            "bytes_alg_per_step": bytes_alg / nsteps,
        },
        # the measured HBM bytes of a step over the same event-timed duration (never above the copy
        # rate of the same run when both cover the same steps)
        "measured_hbm_GBps": measured,
        "kernel_avg_us": {k: 1e3 * v / nsteps for k, v in kms.items()},
        "per_step": {k: (q1[k] - q0[k]) / max(steps, 1)
                     for k in ("frontier", "captures", "solidified", "evaluated", "masks")},
        "copy_GBps": copy_gbps,
        "init_s": init_s,
        "wall_s": wall,
    }
    if not args.no_fast:
        line["fast_mode"] = fast_leg(deck, w0, steps)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(deck, state, w0, steps, after)
        del state, after
    if args.ensemble > 0:
        line["ensemble"] = ensemble_leg(args.ensemble)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
