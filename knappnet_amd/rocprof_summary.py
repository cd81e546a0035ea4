"""Summaries of rocprofv3 result databases for the committed profiles.

    python -m knappnet_amd.rocprof_summary trace <run_results.db>
    python -m knappnet_amd.rocprof_summary traffic <fetch.db> <write.db> <kernel> <out.json>

``trace`` prints per-kernel call counts and durations (a --kernel-trace
--stats run). ``traffic`` reads two --pmc runs (FETCH_SIZE, then WRITE_SIZE;
KiB per dispatch) that each also ran one 1 GiB device-to-device copy
(bench.py --calibrate-copy), scales both counters by what they read for that
copy, and writes the HBM bytes per launch of ``kernel`` as JSON.
"""
from __future__ import annotations

import collections
import json
import re
import sqlite3
import sys

COPY_BYTES = 1 << 30
COPY_KERNEL = "netbc_copy_kernel"


def kernel_name(full: str) -> str:
    m = re.search(r"::(\w+)(<[^>]*>)?\(", full)
    if m:
        return m.group(1) + (m.group(2) or "")
    return full.split("(")[0][:60]


def trace(db: str) -> list[dict]:
    """Per-kernel launches and durations from the dispatch records (ns)."""
    con = sqlite3.connect(db)
    agg = collections.defaultdict(lambda: [0, 0.0])
    for name, start, end in con.execute("select name, start, end from kernels"):
        a = agg[kernel_name(name)]
        a[0] += 1
        a[1] += float(end) - float(start)
    tot = sum(v[1] for v in agg.values()) or 1.0
    rows = [{"kernel": k, "calls": v[0], "total_ns": v[1], "avg_ns": v[1] / v[0],
             "percent": 100.0 * v[1] / tot} for k, v in agg.items()]
    return sorted(rows, key=lambda r: -r["total_ns"])


def counters(db: str, counter: str) -> dict:
    """Per kernel: (dispatches, summed counter value) for one counter."""
    con = sqlite3.connect(db)
    agg = collections.defaultdict(lambda: [0, 0.0])
    seen = set()
    for disp, name, cname, value in con.execute(
            "select dispatch_id, kernel_name, counter_name, value from counters_collection"):
        if cname != counter:
            continue
        k = kernel_name(name)
        agg[k][1] += float(value)
        if (k, disp) not in seen:
            seen.add((k, disp))
            agg[k][0] += 1
    return dict(agg)


def traffic(fetch_db: str, write_db: str, kernel: str) -> dict:
    f = counters(fetch_db, "FETCH_SIZE")
    w = counters(write_db, "WRITE_SIZE")
    if COPY_KERNEL not in f or COPY_KERNEL not in w:
        raise SystemExit("the calibration copy is missing from the runs")
    # the copy kernel reads and writes COPY_BYTES once
    fetch_scale = COPY_BYTES / max(f[COPY_KERNEL][1] * 1024.0, 1.0)
    write_scale = COPY_BYTES / max(w[COPY_KERNEL][1] * 1024.0, 1.0)
    fk, wk = f.get(kernel), w.get(kernel)
    if not fk or not wk:
        raise SystemExit(f"kernel {kernel} not found")
    fb = fk[1] * 1024.0 * fetch_scale
    wb = wk[1] * 1024.0 * write_scale
    return {"kernel": kernel, "launches": fk[0], "fetch_bytes": fb, "write_bytes": wb,
            "bytes_per_launch": (fb + wb) / fk[0], "fetch_scale": fetch_scale,
            "write_scale": write_scale}


def main(argv: list[str]) -> None:
    if argv[:1] == ["trace"]:
        for r in trace(argv[1]):
            print(f"{r['kernel']:36s} calls={r['calls']:7d} total={r['total_ns'] / 1e6:10.2f} ms "
                  f"avg={r['avg_ns'] / 1e3:10.2f} us {r['percent']:6.2f}%")
    elif argv[:1] == ["traffic"]:
        t = traffic(argv[1], argv[2], argv[3])
        t["source"] = argv[5] if len(argv) > 5 else argv[1]
        with open(argv[4], "w") as fh:
            json.dump(t, fh, indent=1)
        print(json.dumps(t, indent=1))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
