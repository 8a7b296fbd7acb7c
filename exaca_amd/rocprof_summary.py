#!/usr/bin/env python3
"""Summaries of rocprofv3 SQLite results for the CA step kernels.

  trace  DB [--last N]            per-kernel dispatch count and mean/min/max duration, over every
                                  dispatch and over each kernel's last N dispatches (the window
                                  that bench.py's per-kernel event timing covers)
  pmc    DB [DB ...]              per-kernel mean of each counter per dispatch (the counter summed
                                  over its instances)
  bytes  --fetch DB --write DB --cal-fetch DB --cal-write DB --copy-bytes B
                                  per-kernel HBM bytes per dispatch: FETCH_SIZE and WRITE_SIZE (KiB)
                                  scaled by copy_bytes / (what the same counter reads for k_copy,
                                  a plain device copy of copy_bytes)

The kernel name is the k_* identifier inside the mangled symbol.
"""
from __future__ import annotations

import argparse
import collections
import re
import sqlite3
import sys


def _kname(symbol: str) -> str:
    m = re.search(r"(k_\w+?)(?:ENS|EPK|Ei|E[A-Z]|$)", symbol)
    if m:
        return m.group(1)
    m = re.search(r"(k_[A-Za-z0-9_]+)", symbol)
    return m.group(1) if m else symbol[:40]


def trace(db: str, last: int) -> None:
    con = sqlite3.connect(db)
    rows = con.execute(
        "select s.kernel_name, d.start, d.end - d.start from rocpd_kernel_dispatch d "
        "join rocpd_info_kernel_symbol s on d.kernel_id = s.id order by d.start")
    per = collections.defaultdict(list)
    for name, start, dur in rows:
        per[_kname(name)].append(dur)
    total = sum(sum(v) for v in per.values())
    print(f"{'kernel':16s} {'calls':>7s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'pct':>6s} "
          f"{'last':>5s} {'avg_us(last)':>12s}")
    for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        w = v[-last:] if last > 0 else v
        print(f"{k:16s} {len(v):7d} {sum(v) / len(v) / 1e3:9.2f} {min(v) / 1e3:9.2f} {max(v) / 1e3:9.2f} "
              f"{100 * sum(v) / total:6.2f} {len(w):5d} {sum(w) / len(w) / 1e3:12.2f}")


def _pmc(dbs: list[str]) -> dict:
    tot = collections.defaultdict(float)
    dur = {}
    for db in dbs:
        con = sqlite3.connect(db)
        for name, cn, v, d, disp in con.execute(
                "select name, counter_name, counter_value, duration, dispatch_id from pmc_events"):
            k = _kname(name)
            tot[(k, cn, db, disp)] += v
            dur[(k, db, disp)] = d
    per = collections.defaultdict(list)
    for (k, cn, db, disp), v in tot.items():
        per[(k, cn)].append(v)
    durs = collections.defaultdict(list)
    for (k, db, disp), d in dur.items():
        durs[k].append(d)
    return {key: (sum(v) / len(v), len(v), sum(durs[key[0]]) / len(durs[key[0]]) / 1e3) for key, v in per.items()}


def pmc(dbs: list[str]) -> None:
    for (k, cn), (avg, n, us) in sorted(_pmc(dbs).items()):
        print(f"{k:16s} {cn:22s} dispatches={n:4d} avg={avg:16.1f}  dur_us={us:8.2f}")


def bytes_per_dispatch(a) -> None:
    f, w = _pmc([a.fetch]), _pmc([a.write])
    cf, cw = _pmc([a.cal_fetch]), _pmc([a.cal_write])
    cal_f = a.copy_bytes / (cf[("k_copy", "FETCH_SIZE")][0] * 1024)
    cal_w = a.copy_bytes / (cw[("k_copy", "WRITE_SIZE")][0] * 1024)
    print(f"calibration: k_copy of {a.copy_bytes} B reads FETCH_SIZE {cf[('k_copy', 'FETCH_SIZE')][0]:.0f} KiB, "
          f"WRITE_SIZE {cw[('k_copy', 'WRITE_SIZE')][0]:.0f} KiB -> factors {cal_f:.3f} (fetch), {cal_w:.3f} (write)")
    print(f"{'kernel':16s} {'dispatches':>10s} {'fetch_MB':>10s} {'write_MB':>10s} {'total_MB':>10s} {'dur_us':>8s}")
    total = 0.0
    for (k, cn), (avg, n, us) in sorted(f.items()):
        if cn != "FETCH_SIZE" or k == "k_copy":
            continue
        fb = avg * 1024 * cal_f
        wb = w.get((k, "WRITE_SIZE"), (0.0, 0, 0.0))[0] * 1024 * cal_w
        total += fb + wb
        print(f"{k:16s} {n:10d} {fb / 1e6:10.3f} {wb / 1e6:10.3f} {(fb + wb) / 1e6:10.3f} {us:8.2f}")
    print(f"{'sum':16s} {'':10s} {'':10s} {'':10s} {total / 1e6:10.3f}")


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("--last", type=int, default=0)
    p = sub.add_parser("pmc")
    p.add_argument("db", nargs="+")
    b = sub.add_parser("bytes")
    for o in ("--fetch", "--write", "--cal-fetch", "--cal-write"):
        b.add_argument(o, required=True)
    b.add_argument("--copy-bytes", type=int, required=True)
    a = ap.parse_args()
    if a.cmd == "trace":
        trace(a.db, a.last)
    elif a.cmd == "pmc":
        pmc(a.db)
    else:
        bytes_per_dispatch(a)
    return 0


if __name__ == "__main__":
    sys.exit(main())
