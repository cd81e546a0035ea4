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
  sources;
- many small graphs (``netbc_betweenness_batch``): 1,024 Watts-Strogatz
  graphs, n uniform in [50, 2000], k = 6, p = 0.1.

The search is coordinate descent from the compiled-in values, one setting
at a time in the order of ``KNOBS``: slots per super-batch, the direction
factor alpha, the sweep group size, the sweep's and the search step's
workgroup sizes, the hub threshold and hub task size (breadth-first
classes); slots and the bucket width (integer lengths: the candidates are
the widths up to the shortest arc, the only ones that keep every bucket
settled at once); the workgroup size (many small graphs). Each candidate is
timed ``repeats`` times (wall time of one call, the median kept) and
replaces the current value only if it is at least 2% faster. Rebuild the
library after writing the table.
"""
from __future__ import annotations

import argparse
import statistics
import time
from pathlib import Path

import numpy as np

from . import generate as gen
from .engine import Graph, betweenness_batch, device_info

PKG = Path(__file__).resolve().parent

CANDIDATES = {
    "slots": [16, 32, 64, 128],
    "alpha": [0.25, 0.5, 1.0, 2.0, 4.0],
    "sweep_group": [4, 8, 12, 16, 32, 64],
    "sweep_block": [128, 256, 512, 1024],
    "step_block": [128, 256, 512, 1024],
    "hub_in": [32, 64, 128, 256],
    "hub_chunk": [256, 512, 1024, 2048],
    "bucket_width": [1, 2, 4, 8, 16],
}
# compiled-in starting values (tuning.h's defaults when no class matches)
START = {"slots": 64, "alpha": 1.0, "sweep_group": 32, "sweep_block": 256, "step_block": 256,
         "hub_in": 64, "hub_chunk": 512, "bucket_width": 1}
KNOBS = {
    0: ["slots", "alpha", "sweep_group", "sweep_block", "step_block", "hub_in", "hub_chunk"],
    1: ["slots", "bucket_width"],
    2: ["sweep_block"],
}
SMALL_BLOCKS = [256, 512, 1024]


def workloads(quick: bool):
    """(kind, directed, description, payload, weighted): payload is
    (edge list, sources) for one-graph classes and the graph list for the
    many-graphs class."""
    el = gen.rmat(18, 16, directed=True, seed=1)
    yield (0, 1, "R-MAT scale 18 directed, first 32,768 sources",
           (el, np.arange(8192 if quick else 32768)), False)
    if not quick:
        el = gen.rmat(22, 16, directed=False, seed=1)
        src = gen.sample_sources(el.n, 8192, 7)[:4096]
        yield (0, 0, "R-MAT scale 22 undirected, 4,096 sampled sources", (el, src), False)
    el = gen.ba_weighted()
    yield (1, -1, "BA n=200,000 integer weights 1-16, first 8,192 sources",
           (el, np.arange(2048 if quick else 8192)), True)
    els = gen.ws_batch(256 if quick else 1024, 11)
    yield (2, 0, f"{len(els):,} Watts-Strogatz graphs, n 50-2,000, k 6, p 0.1",
           [(e.n, e.src, e.tgt) for e in els], False)


def candidates(knob: str, shortest: int) -> list:
    if knob == "bucket_width":
        return [w for w in CANDIDATES[knob] if w <= max(shortest, 1)]
    return CANDIDATES[knob]


class Timer:
    """Times one call of a class's workload under a setting."""

    def __init__(self, kind: int, payload, weighted: bool, repeats: int):
        self.kind, self.weighted, self.repeats = kind, weighted, repeats
        if kind == 2:
            self.graphs = payload
            self.shortest = 1
            betweenness_batch(self.graphs[:16])  # warm up
            return
        el, self.sources = payload
        self.g = Graph(el.n, el.src, el.tgt, el.weights, el.directed)
        self.shortest = int(el.weights.min()) if weighted else 1
        self.g.betweenness_sources(self.sources[:1024], weighted=weighted)  # CSR, workspace

    def __call__(self, s: dict) -> float:
        ts = []
        for _ in range(self.repeats):
            t = time.perf_counter()
            if self.kind == 2:
                betweenness_batch(self.graphs, sweep_block=s["sweep_block"])
            else:
                self.g.betweenness_sources(self.sources, weighted=self.weighted, **s)
            ts.append(time.perf_counter() - t)
        return statistics.median(ts)


def tune_class(timer: Timer, kind: int, repeats: int, log):
    best = dict(START)
    if kind == 1:
        best["slots"] = 16
    if kind == 2:
        best["sweep_block"] = 512
    best_t = timer(best)
    log(f"  start {best}: {best_t:.3f} s")
    for knob in KNOBS[kind]:
        cands = SMALL_BLOCKS if kind == 2 else candidates(knob, timer.shortest)
        for v in cands:
            if v == best[knob]:
                continue
            trial = dict(best, **{knob: v})
            t = timer(trial)
            log(f"  {knob}={v}: {t:.3f} s")
            if t < best_t * 0.98:
                best, best_t = trial, t
        log(f"  -> {knob}={best[knob]} ({best_t:.3f} s)")
    return best, best_t


def write_table(path: Path, entries, device: str) -> None:
    rows = []
    for kind, directed, desc, s in entries:
        slots = 0 if kind == 2 else s["slots"]
        rows.append(f'    {{{kind}, {directed}, 0, 40, 0.0, 1e30, {slots}, {s["alpha"]!r}, '
                    f'{s["sweep_group"]}, {s["sweep_block"]}, {s["step_block"]}, {s["hub_in"]}, '
                    f'{s["hub_chunk"]}, {s["bucket_width"]}, "{desc}"}},')
    path.write_text(f"""// Engine settings per graph class, written by knappnet_amd/autotune.py from
// timed runs on {device} (the search, candidates and timings it used are
// kept with the profiles). Regenerate on a new device with
// `python -m knappnet_amd.autotune`.
#pragma once

namespace netbc {{

struct TuneEntry {{
    int kind;           // 0 breadth-first, 1 integer lengths (Dial), 2 many small graphs
                        // (netbc_betweenness_batch: only sweep_block, the workgroup size)
    int directed;       // 0 undirected, 1 directed, -1 either
    int log2n_min;      // node-count class: floor(log2 n) in [min, max]
    int log2n_max;
    double deg_min;     // stored arcs per node in [min, max)
    double deg_max;
    int slots;          // 64-source batches per super-batch
    double alpha;       // bottom-up when frontier out-arcs * alpha > open in-arcs
    int sweep_group;    // entries per dependency-sweep task at most
    int sweep_block;    // threads per workgroup of the dependency sweep
    int step_block;     // threads per workgroup of the bottom-up step and hub tasks
    int hub_in;         // in-degree above which a node is a hub
    int hub_chunk;      // in-arcs per hub task
    int bucket_width;   // integer lengths: distances per bucket
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
    for kind, directed, desc, payload, weighted in workloads(args.quick):
        log(desc)
        timer = Timer(kind, payload, weighted, args.repeats)
        best, _ = tune_class(timer, kind, args.repeats, log)
        entries.append((kind, directed, desc, best))
        del timer
    # breadth-first classes first: the library takes the first entry that
    # matches a graph
    write_table(Path(args.out), entries, device.strip(" ,") or "the device")
    log(f"wrote {args.out}")
    if args.report:
        Path(args.report).write_text("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
