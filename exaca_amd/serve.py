"""Serving loop for process-map solves on one GPU (continuous batching over an Ensemble).

Requests arrive over time; each is initialised on a host thread (exaca_prepare), admitted into a
free ensemble slot when device memory allows, stepped together with every other resident solve in
batched launches, and released as soon as its own termination check ends it. A sequential loop of
the plain path over the same requests is the reference point for throughput; both report a digest
of every result so that identical outputs can be checked.
"""
from __future__ import annotations

import collections
import concurrent.futures as cf
import hashlib
import time

import numpy as np

from . import FIELD_CELLTYPE, FIELD_GRAINID, Ensemble, Prepared, Solver


def digest(cycle: int, ct: np.ndarray, gid: np.ndarray) -> str:
    h = hashlib.sha1()
    h.update(np.int64(cycle).tobytes())
    h.update(np.ascontiguousarray(ct, dtype=np.int32).tobytes())
    h.update(np.ascontiguousarray(gid, dtype=np.int32).tobytes())
    return h.hexdigest()


def run_sequential(requests, realtime: bool = True, log=None) -> dict:
    """The plain GPU path, one request after another in arrival order (host init + run + download
    each). With realtime a request is not started before its arrival time; its latency is its
    completion time minus its arrival."""
    order = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    out = {}
    finished = {}
    gpu_ms = 0.0
    init_s = 0.0
    for r in order:
        if realtime:
            wait = r.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        ti = time.perf_counter()
        with Solver(r.deck, seed=r.seed) as sol:
            init_s += time.perf_counter() - ti
            cycle, ms = sol.run()
            gpu_ms += ms
            out[r.index] = digest(cycle, sol.field(FIELD_CELLTYPE), sol.field(FIELD_GRAINID))
        finished[r.index] = time.perf_counter() - t0
        if log:
            log(f"sequential {r.index}: {r.nx}x{r.ny}x{r.nz} cycle {cycle} {ms:.1f} ms")
    wall = time.perf_counter() - t0
    lat = np.array([finished[r.index] - (r.arrival if realtime else 0.0) for r in requests])
    first = min(r.arrival for r in requests) if realtime else 0.0
    return {
        "wall_s": wall,
        "gpu_s": gpu_ms / 1e3,
        "init_s": init_s,
        "digests": out,
        "solves_per_s": len(requests) / (max(finished.values()) - first),
        "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
    }


def run_served(requests, *, slots: int = 256, arena_bytes: int = 200 << 30, chunk: int = 50,
               host_threads: int = 8, max_prepared: int = 48, autotune: bool = True, tune_at: int | None = None,
               realtime: bool = True, max_steps: int = 5_000_000, log=None) -> dict:
    """Continuous batching. With realtime, request i becomes visible at its arrival time;
    otherwise all are visible at once (the saturated ensemble). Returns latencies (arrival to
    completion), makespan and per-request digests.

    max_prepared bounds the requests initialised on the host and not yet admitted (their host
    state waits in memory); admitted requests do not count, so every slot can be resident. The
    autotuner runs once, when tune_at domains are resident (default: a quarter of min(n, slots)).
    A domain stopped by a device error is a failed request: its digest is "error:<flags>" and its
    index is listed under "failed". More than max_steps steps in all raise RuntimeError."""
    n = len(requests)
    order = sorted(requests, key=lambda r: r.arrival)
    pool = cf.ThreadPoolExecutor(max_workers=host_threads)
    futures = {}
    cells = {}
    waiting = collections.deque()
    running = {}
    finished = {}
    digests = {}
    failed = []
    admitted_at = {}
    tuned = None
    if tune_at is None:
        tune_at = max(1, min(n, slots) // 4)
    nxt = 0
    steps = 0
    peak_resident = 0
    with Ensemble(slots, arena_bytes) as ens:
        t0 = time.perf_counter()
        while len(finished) < n:
            now = time.perf_counter() - t0
            while nxt < n and (not realtime or order[nxt].arrival <= now) and len(waiting) < max_prepared:
                r = order[nxt]
                futures[r.index] = pool.submit(Prepared, r.deck, r.seed)
                waiting.append(r)
                nxt += 1
            while waiting and futures[waiting[0].index].done():
                r = waiting[0]
                prep = futures[r.index].result()
                slot = ens.admit(prep)
                if slot is None:
                    break
                waiting.popleft()
                running[slot] = r
                admitted_at[r.index] = time.perf_counter() - t0
                # the device holds the state now: only the cell count is kept
                cells[r.index] = prep.cells
                prep.close()
                del futures[r.index]
            peak_resident = max(peak_resident, len(running))
            if not running:
                time.sleep(0.0005)
                continue
            if steps > max_steps:
                raise RuntimeError(f"served ensemble exceeded {max_steps} steps with {len(running)} domains running")
            if autotune and tuned is None and len(running) >= tune_at:
                tuned = ens.autotune(1)
                steps += 5 * chunk
            ens.advance(chunk)
            steps += chunk
            st, err = ens.poll_errors()
            done_now = time.perf_counter() - t0
            for slot in [s for s in running if st[s] > 0 or err[s] != 0]:
                r = running.pop(slot)
                if err[slot]:
                    digests[r.index] = f"error:{int(err[slot])}"
                    failed.append(r.index)
                else:
                    ncell = cells.pop(r.index)
                    ct = ens.field(slot, FIELD_CELLTYPE, ncell)
                    gid = ens.field(slot, FIELD_GRAINID, ncell)
                    digests[r.index] = digest(int(st[slot]), ct, gid)
                finished[r.index] = done_now
                ens.release(slot)
                if log:
                    log(f"served {r.index}: cycle {st[slot]} error {int(err[slot])} latency {done_now - r.arrival:.3f} s")
        makespan = time.perf_counter() - t0
    pool.shutdown()
    lat = np.array([finished[r.index] - (r.arrival if realtime else 0.0) for r in requests])
    first = min(r.arrival for r in requests) if realtime else 0.0
    return {
        "makespan_s": makespan,
        "solves_per_s": n / (max(finished.values()) - first),
        "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
        "mean_s": float(lat.mean()),
        "blocks_per_slot": tuned,
        "steps": steps,
        "peak_resident": peak_resident,
        "failed": failed,
        "digests": digests,
    }
