"""Chooses the engine's settings per graph class by timing candidate values
on the current device, and writes them as the table the library compiles in
(``knappnet_amd/csrc/tuning.h``) together with a report of every timing.

    python -m knappnet_amd.autotune [--out PATH] [--report PATH] [--quick]

Each class is represented by one seeded workload, timed on a prefix of its
sources (enough for several super-batches):

- breadth-first, directed: R-MAT scale 18 (edge factor 16), 32,768 sources;
- breadth-first, undirected: R-MAT scale 22, 4,096 of its 8,192 sampled
  sources;
- integer lengths (Dial): Barabasi-Albert n = 200,000, weights 1-16, 8,192
  sources.

The search is coordinate descent from the compiled-in values: slots per
super-batch, then the direction factor alpha, then the sweep group size (the
last two only for breadth-first classes), each candidate timed ``repeats``
times (wall time of ``netbc_betweenness_sources``, the median kept). A
candidate replaces the current value only if it is at least 2% faster.
Rebuild the library after writing the table.
"""
from __future__ import annotations

import argparse
import statistics
import time
from pathlib import Path

import numpy as np

from . import generate as gen
from .engine import Graph, device_info

PKG = Path(__file__).resolve().parent

CANDIDATES = {
    "slots": [16, 32, 64, 128],
    "alpha": [0.25, 0.5, 1.0, 2.0, 4.0],
    "sweep_group": [4, 8, 12, 16, 32, 64],
}
START = {"slots": 64, "alpha": 1.0, "sweep_group": 32}


def workloads(quick: bool):
    """(class key, description, edge list, sources, weighted)."""
    el = gen.rmat(18, 16, directed=True, seed=1)
    yield ((0, 1), "R-MAT scale 18 directed, first 32,768 sources", el,
           np.arange(8192 if quick else 32768), False)
    if not quick:
        el = gen.rmat(22, 16, directed=False, seed=1)
        src = gen.sample_sources(el.n, 8192, 7)[:4096]
        yield ((0, 0), "R-MAT scale 22 undirected, 4,096 sampled sources", el, src, False)
    el = gen.ba_weighted()
    yield ((1, -1), "BA n=200,000 integer weights 1-16, first 8,192 sources", el,
           np.arange(2048 if quick else 8192), True)


def time_run(g: Graph, sources, weighted: bool, setting: dict, repeats: int) -> float:
    ts = []
    for _ in range(repeats):
        t = time.perf_counter()
        g.betweenness_sources(sources, weighted=weighted, slots=setting["slots"],
                              alpha=setting["alpha"], sweep_group=setting["sweep_group"])
        ts.append(time.perf_counter() - t)
    return statistics.median(ts)


def tune_class(g: Graph, sources, weighted: bool, knobs, repeats: int, log):
    best = dict(START)
    if weighted:
        best["slots"] = 16
    g.betweenness_sources(sources[:1024], weighted=weighted)  # warm up: CSR, workspace
    best_t = time_run(g, sources, weighted, best, repeats)
    log(f"  start {best}: {best_t:.3f} s")
    for knob in knobs:
        for v in CANDIDATES[knob]:
            if v == best[knob]:
                continue
            trial = dict(best, **{knob: v})
            t = time_run(g, sources, weighted, trial, repeats)
            log(f"  {knob}={v}: {t:.3f} s")
            if t < best_t * 0.98:
                best, best_t = trial, t
        log(f"  -> {knob}={best[knob]} ({best_t:.3f} s)")
    return best, best_t


def write_table(path: Path, entries, device: str) -> None:
    rows = []
    for (kind, directed), desc, s in entries:
        rows.append(f'    {{{kind}, {directed}, 0, 40, 0.0, 1e30, {s["slots"]}, {s["alpha"]!r}, '
                    f'{s["sweep_group"]}, "{desc}"}},')
    path.write_text(f"""// Engine settings per graph class, written by knappnet_amd/autotune.py from
// timed runs on {device} (the search, candidates and timings it used are
// kept with the profiles). Regenerate on a new device with
// `python -m knappnet_amd.autotune`.
#pragma once

namespace netbc {{

struct TuneEntry {{
    int kind;           // path lengths: 0 breadth-first, 1 integer (Dial)
    int directed;       // 0 undirected, 1 directed, -1 either
    int log2n_min;      // node-count class: floor(log2 n) in [min, max]
    int log2n_max;
    double deg_min;     // stored arcs per node in [min, max)
    double deg_max;
    int slots;          // 64-source batches per super-batch
    double alpha;       // bottom-up when frontier out-arcs * alpha > open in-arcs
    int sweep_group;    // entries per dependency-sweep task at most
    const char* tuned_on;
}};

constexpr TuneEntry kTuning[] = {{
{chr(10).join(rows)}
}};

}}  // namespace netbc
""")


def main(argv=None) -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=str(PKG / "csrc" / "tuning.h"))
    ap.add_argument("--report", default=None)
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--quick", action="store_true", help="small source prefixes (a smoke run)")
    args = ap.parse_args(argv)
    lines = []

    def log(s):
        print(s, flush=True)
        lines.append(s)

    device = device_info()
    log(f"# autotune on {device}")
    entries = []
    for key, desc, el, sources, weighted in workloads(args.quick):
        log(f"{desc} (n={el.n}, m={el.m})")
        g = Graph(el.n, el.src, el.tgt, el.weights, el.directed)
        knobs = ["slots"] if weighted else ["slots", "alpha", "sweep_group"]
        best, t = tune_class(g, np.asarray(sources), weighted, knobs, args.repeats, log)
        entries.append((key, desc, best))
        del g
    # breadth-first classes first: the library takes the first entry that
    # matches a graph
    write_table(Path(args.out), entries, device.strip(" ,") or "the device")
    log(f"wrote {args.out}")
    if args.report:
        Path(args.report).write_text("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
