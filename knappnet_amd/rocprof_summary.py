"""Summaries of rocprofv3 result databases for the committed profiles.

    python -m knappnet_amd.rocprof_summary trace <run_results.db>
    python -m knappnet_amd.rocprof_summary traffic <fetch.db> <write.db> <kernel> <out.json>
    python -m knappnet_amd.rocprof_summary counters <run_results.db> <counter>...
    python -m knappnet_amd.rocprof_summary gather <fetch.db>

``trace`` prints per-kernel call counts and durations (a --kernel-trace
--stats run). ``traffic`` reads two --pmc runs (FETCH_SIZE, then WRITE_SIZE;
KiB per dispatch) that each also ran one 1 GiB device-to-device copy
(bench.py --calibrate-copy), scales both counters by what they read for that
copy, and writes the HBM bytes per launch of ``kernel`` as JSON.
``counters`` prints, per kernel, the mean per dispatch of each counter of a
--pmc run, and the L2 hit rate when TCC_HIT_sum and TCC_MISS_sum are among
them. ``gather`` prints what FETCH_SIZE reads for the calibration gather
(bench.py --calibrate-copy: one random 8-byte read per 128-byte line of a
1 GiB buffer) per read, raw and with the copy's scale.
"""
from __future__ import annotations

import collections
import json
import re
import sqlite3
import sys

COPY_BYTES = 1 << 30
COPY_KERNEL = "netbc_copy_kernel"
GATHER_KERNEL = "netbc_gather_kernel"
GATHER_LINE = 128


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


def counter_table(db: str, names: list[str]) -> list[dict]:
    """Per kernel: dispatches and the mean per dispatch of each counter."""
    per = {c: counters(db, c) for c in names}
    kernels = sorted({k for c in per.values() for k in c},
                     key=lambda k: -max(per[c].get(k, (0, 0.0))[1] for c in names))
    rows = []
    for k in kernels:
        r = {"kernel": k, "dispatches": max(per[c].get(k, (0, 0.0))[0] for c in names)}
        for c in names:
            d, v = per[c].get(k, (0, 0.0))
            r[c] = v / d if d else None
        if r.get("TCC_HIT_sum") is not None and r.get("TCC_MISS_sum") is not None:
            tot = r["TCC_HIT_sum"] + r["TCC_MISS_sum"]
            r["l2_hit_pct"] = 100.0 * r["TCC_HIT_sum"] / tot if tot else None
        rows.append(r)
    return rows


def gather_calibration(fetch_db: str) -> dict:
    f = counters(fetch_db, "FETCH_SIZE")
    if COPY_KERNEL not in f or GATHER_KERNEL not in f:
        raise SystemExit("the calibration copy or gather is missing from the run")
    scale = COPY_BYTES / max(f[COPY_KERNEL][1] * 1024.0, 1.0)
    reads = COPY_BYTES // GATHER_LINE
    raw = f[GATHER_KERNEL][1] * 1024.0 / reads
    return {"reads": reads, "line": GATHER_LINE, "raw_bytes_per_read": raw,
            "copy_fetch_scale": scale, "scaled_bytes_per_read": raw * scale}


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
    elif argv[:1] == ["counters"]:
        names = argv[2:]
        for r in counter_table(argv[1], names):
            vals = " ".join(f"{c}={r[c]:.4g}" if r[c] is not None else f"{c}=-" for c in names)
            hit = f" l2_hit={r['l2_hit_pct']:.1f}%" if r.get("l2_hit_pct") is not None else ""
            print(f"{r['kernel']:44s} dispatches={r['dispatches']:6d} {vals}{hit}")
    elif argv[:1] == ["gather"]:
        print(json.dumps(gather_calibration(argv[1]), indent=1))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
