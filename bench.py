"""Benchmark: exact betweenness GTEPS (sources x edges / s) on one MI355X.

Default workload: directed R-MAT scale 18 (n = 262,144, edge factor 16,
a,b,c,d = 0.57,0.19,0.19,0.05, seed 1, loops dropped, duplicates removed,
isolates kept; m = 3,939,372), every node a source, FP64 sigma, dependencies
and scores. One step = one full all-sources betweenness through the C ABI
(``netbc_betweenness_sources``), with the graph already resident on the
device (the warm-up call builds the device CSR).

The printed JSON line carries:
- ``roofline`` for the dominant kernel (the dependency sweep), from its
  launch durations measured with HIP events on the stream it runs on and
  its algorithmic bytes (8 B per reached out-arc: column index and head
  level; per reached node 32 B, row offsets, sigma read, coefficient and
  dependency writes, or 24 B in runs of 4,096 or more sources, whose
  dependencies go into one sum per (slot, node) instead of a row; summed
  over sources);
- ``cpu_baseline``: the CPU oracle (the reference's algorithm restated in
  C, all host cores) timed on a strided sample of the same sources, its
  time extrapolated to every source by reached arcs, and a check of the
  device result on that sample;
- ``configs``: one timed run of each other workload (G(2000, 10000) from
  seed 0; Barabasi-Albert n = 200,000 with integer weights 1-16; 4,096
  Watts-Strogatz graphs in one call; undirected R-MAT scale 22 with 8,192
  sources from seed 7, FP64 and FP32 dependencies), each with its own
  roofline, CPU baseline and check (``--no-configs`` skips them).

``--edge`` times edge betweenness (``netbc_edge_betweenness_sources``)
on the same workload instead; its roofline keeps the node sweep's byte
model, a lower bound for the edge sweep (which also updates one 8-byte
credit per slot and arc it meets). ``--subset`` times betweenness for
sources and targets (``netbc_betweenness_subset_ex``: the workload's
sources, targets every third node, NetworkX's subset rescaling, node
values only); its CPU baseline is the oracle's single-threaded
restatement. ``--closeness`` times normalized closeness of every node
(``netbc_closeness_ex``, unreachable nodes ignored: reached / sum of
distances) on the same workload; the roofline kernel is the bottom-up step
of the reachability-only search (``bfs_pull_kernel<false, false>``), whose
algorithmic bytes the engine counts as it runs (8 per pair examined, 8 per
in-row looked up, 12 per in-arc read, 40 per pair reached, 24 per stale
level entry cleared), and the CPU baseline is the oracle's distance
profiles of a strided sample of sources on every host core.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("GTEPS for exact betweenness (sources×edges/s), FP32-δ vs FP64 at rel err "
          "≤1e-6")
METRIC_EDGE = "GTEPS for exact edge betweenness (sources×edges/s)"
METRIC_SUBSET = "GTEPS for exact betweenness for sources and targets (sources×edges/s)"
METRIC_CLOSENESS = "GTEPS for closeness centrality (sources×edges/s)"
HBM_PEAK_GBS = 8000.0


def build_workload(name: str):
    from knappnet_amd import generate as gen
    if name == "rmat18":
        el = gen.rmat(18, 16, directed=True, seed=1)
        return el, None, dict(workload="rmat18_directed_all_sources", scale=18, edge_factor=16,
                              abcd=[0.57, 0.19, 0.19, 0.05], seed=1)
    if name == "er2000":
        el = gen.er_2000()
        return el, None, dict(workload="er_gnm_2000_10000_all_sources", seed=0)
    if name == "ba200k":
        el = gen.ba_weighted()
        return el, None, dict(workload="ba_200000_m4_int_weights_1_16_all_sources", seed=3)
    if name == "ba200k-real":
        el = gen.ba_real()
        return el, None, dict(workload="ba_200000_m4_real_weights_1_16_all_sources", seed=3)
    if name == "rmat22":
        el = gen.rmat(22, 16, directed=False, seed=1)
        src = gen.sample_sources(el.n, 8192, 7)
        return el, src, dict(workload="rmat22_undirected_8192_sampled_sources", scale=22,
                             edge_factor=16, seed=1, source_seed=7)
    raise SystemExit(f"unknown workload {name}")


def med(run, steps):
    """The last result of `steps` runs and their median wall time (s)."""
    ts, last = [], None
    for _ in range(steps):
        t = time.perf_counter()
        last = run()
        ts.append(time.perf_counter() - t)
    return last, float(np.median(ts))


def plan_legs(els, graphs, parts) -> dict:
    """Repeated batches through netbc_batch_plan_*: the 4,096-graph batch
    replayed from a captured HIP graph and launched kernel by kernel
    (seconds per batch, bit-identity with the batch entry's scores), and
    the conditional-uniform-graph pattern of one call per replicate
    (r/R/cug.R:42-47): 64 Watts-Strogatz replicates of one shape (n = 1000,
    k = 6, so m = 3000) through the batch entry one graph per call, a
    one-graph plan captured and uncaptured, and all 64 in one captured
    plan (graphs per second each)."""
    from knappnet_amd import generate as gen
    from knappnet_amd.engine import BatchPlan, betweenness_batch
    out = {}
    ns = [e.n for e in els]
    ms = [e.m for e in els]
    for capture in (True, False):
        t = time.perf_counter()
        plan = BatchPlan(ns, ms, capture=capture)
        made = time.perf_counter() - t
        got, _ = plan.run(graphs)
        (res, _), sec = med(lambda: plan.run(graphs), 2)
        same = all(np.array_equal(a, b) for a, b in zip(res, parts)) and \
            all(np.array_equal(a, b) for a, b in zip(got, parts))
        out["captured" if capture else "uncaptured"] = {
            "seconds": round(sec, 3), "graphs_per_s": round(len(els) / sec, 1),
            "plan_setup_s": round(made, 3), "bit_identical_to_batch_call": same}
        plan.close()
    reps = [gen.small_world(1000, 6, 0.1, 5000 + i, 0) for i in range(64)]
    rg = [(e.n, e.src, e.tgt) for e in reps]
    want, _ = betweenness_batch(rg)
    legs = {}
    t = time.perf_counter()
    one = [betweenness_batch([g])[0][0] for g in rg]
    legs["batch_call_per_replicate"] = time.perf_counter() - t
    ok = all(np.array_equal(a, b) for a, b in zip(one, want))
    for capture in (True, False):
        plan = BatchPlan([1000], [3000], capture=capture)
        plan.run([rg[0]])
        t = time.perf_counter()
        one = [plan.run([g])[0][0] for g in rg]
        legs[("captured" if capture else "uncaptured") + "_plan_per_replicate"] = \
            time.perf_counter() - t
        ok = ok and all(np.array_equal(a, b) for a, b in zip(one, want))
        plan.close()
    plan = BatchPlan([1000] * 64, [3000] * 64)
    plan.run(rg)
    t = time.perf_counter()
    res, _ = plan.run(rg)
    legs["captured_plan_all_64"] = time.perf_counter() - t
    ok = ok and all(np.array_equal(a, b) for a, b in zip(res, want))
    plan.close()
    out["replicates_n1000_m3000"] = {
        **{k: {"seconds": round(v, 4), "graphs_per_s": round(64 / v, 1)} for k, v in legs.items()},
        "bit_identical": ok}
    print("# plans: " + ", ".join(f"{k} {v['graphs_per_s']}/s" for k, v in
                                  out["replicates_n1000_m3000"].items() if isinstance(v, dict)),
          flush=True)
    return out


def run_ws_batch(args) -> None:
    """4,096 Watts-Strogatz graphs (n uniform in [50, 2000], k = 6, p = 0.1)
    in one netbc_betweenness_batch call, against one net_betweenness call
    per graph (timed on a prefix of the graphs)."""
    from knappnet_amd import generate as gen
    from knappnet_amd.engine import Graph, betweenness_batch, device_info
    t0 = time.perf_counter()
    els = gen.ws_batch(4096, 11)
    gen_s = time.perf_counter() - t0
    graphs = [(e.n, e.src, e.tgt) for e in els]
    work = sum(e.n * e.m for e in els)  # sources x edges
    print(f"# device: {device_info()}; 4096 Watts-Strogatz graphs, sum n = "
          f"{sum(e.n for e in els)}, sum m = {sum(e.m for e in els)} "
          f"(generated in {gen_s:.1f}s)", flush=True)
    for i in range(args.warmup):
        parts, diag = betweenness_batch(graphs, precision=args.precision)
    times = []
    for i in range(args.steps):
        t = time.perf_counter()
        parts, diag = betweenness_batch(graphs, precision=args.precision)
        times.append(time.perf_counter() - t)
        print(f"# step {i}: {times[-1]:.3f}s kernel_ms={diag['kernel_ms']:.1f}", flush=True)
    sec = sum(times) / len(times)
    # one launch per graph, the reference's caller pattern (r/R/cug.R:42-47)
    k1 = min(len(els), 256)
    t = time.perf_counter()
    one = [Graph(e.n, e.src, e.tgt, None, False).betweenness() for e in els[:k1]]
    t1 = time.perf_counter() - t
    per_graph_s = t1 / sum(e.n * e.m for e in els[:k1]) * work
    same = all(np.array_equal(a, b) for a, b in zip(one, parts[:k1]))

    cpu = None
    if args.cpu_seconds > 0:
        from oracle.oracle import OracleGraph, available_cores
        step = max(1, len(els) // 64)
        sel = list(range(0, len(els), step))
        t = time.perf_counter()
        want = [OracleGraph(els[i].n, els[i].src, els[i].tgt, None, False).betweenness(threads=0)
                for i in sel]
        ct = time.perf_counter() - t
        differ = [i for i, w in zip(sel, want) if not np.array_equal(parts[i], w)]
        cw = sum(els[i].n * els[i].m for i in sel)
        cpu = {"value": round(cw / ct / 1e9, 4), "unit": "GTEPS", "cores": available_cores(),
               "kind": "port", "seconds": round(ct, 3),
               "sample": f"{len(sel)} of 4096 graphs (every {step}th), one at a time, all host "
                         f"threads, CPU: {cpu_model()}",
               "check": "bit-exact" if not differ else f"{len(differ)} graphs differ"}
    line = {
        "metric": METRIC, "value": round(work / sec / 1e9, 3), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f32-delta/f64-sigma", "data": "synthetic",
        "config": {"workload": "watts_strogatz_4096_graphs_all_sources", "k": 6, "p": 0.1,
                   "n_range": [50, 2000], "base_seed": 11, "graphs_per_s": round(4096 / sec, 1)},
        "one_graph_per_launch": {"seconds_extrapolated": round(per_graph_s, 3),
                                 "timed_graphs": k1, "speedup": round(per_graph_s / sec, 2),
                                 "bit_identical": same},
        "plans": plan_legs(els, graphs, parts),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def sweep_node_bytes(sources: int, exact: bool = False) -> int:
    """Algorithmic bytes per reached node of the breadth-first sweep: row
    offsets 8, sigma read 8, coefficient write 8, and the dependency write 8
    unless the run's sources are batched freely (4,096 or more, not exact
    order), where each entry's dependencies are added into one word per
    (slot, node) and the row is not stored."""
    return 24 if sources >= 4096 and not exact else 32


def sweep_roofline(kernel: str, diag: dict, nbytes: float) -> dict:
    """roofline of a run's timed kernel: `nbytes` algorithmic bytes over the
    summed durations of its launches (HIP events on the engine stream)."""
    sec = diag["dominant_ms"] / 1e3
    achieved = nbytes / sec / 1e9 if sec > 0 else 0.0
    return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None, "launches": diag["dominant_launches"],
            "avg_launch_us": round(diag["dominant_ms"] * 1e3 / max(diag["dominant_launches"], 1),
                                   2),
            "algorithmic_bytes_per_run": int(nbytes)}


def cpu_leg(orun, drun, pool, m: int, k: int, full_arcs: int, target_s: float):
    """The oracle on a strided sample of `pool` (grown until it takes about
    target_s / 4), checked against the device on the same sample; the full
    run's CPU time is extrapolated by reached arcs (the device counts them
    for the sample and for the full run)."""
    from oracle.oracle import available_cores
    take = min(len(pool), 16)
    while True:
        stride = max(1, len(pool) // take)
        sample = np.asarray(pool)[::stride][:take]
        t = time.perf_counter()
        want = orun(sample)
        ct = time.perf_counter() - t
        if ct >= target_s / 4 or take >= len(pool):
            break
        take = min(len(pool), int(take * max(2.0, target_s / 2 / max(ct, 1e-3))))
    got, d = drun(sample)
    exact = bool(np.array_equal(got, want))
    nz = (got != 0) | (want != 0)
    rel = float(np.max(np.abs(got - want)[nz] / np.maximum(np.abs(want[nz]), 1e-300))) \
        if nz.any() else 0.0
    full_s = ct * full_arcs / max(d["reached_arcs"], 1)
    return {"value": round(k * m / full_s / 1e9, 4), "unit": "GTEPS",
            "cores": available_cores(), "kind": "port", "seconds": round(ct, 3),
            "sample": f"{len(sample)} of {k} sources (every {stride}th), all host threads, "
                      f"CPU time extrapolated to every source by reached arcs, CPU: "
                      f"{cpu_model()}",
            "check": "bit-exact" if exact else f"max rel err {rel:.3e}"}


def extra_configs(cpu_seconds: float) -> dict:
    """One timed run of each other workload, with its own roofline, CPU
    baseline and parity check: G(2000, 10000) from seed 0 (all sources),
    Barabasi-Albert n = 200,000 with integer weights 1-16 (all sources),
    4,096 Watts-Strogatz graphs in one call, and undirected R-MAT scale 22
    with the 8,192 sources drawn from seed 7 (FP64, and FP32 dependencies
    against it)."""
    from knappnet_amd import generate as gen
    from knappnet_amd.engine import Graph, betweenness_batch
    from oracle.oracle import OracleGraph, available_cores
    out = {}
    target = max(cpu_seconds / 3, 1.0) if cpu_seconds > 0 else 0.0

    # G(2000, 10000), every node a source (the oracle runs it whole)
    el = gen.er_2000()
    g = Graph(el.n, el.src, el.tgt, None, False)
    g.betweenness_sources()
    (res, d), sec = med(lambda: g.betweenness_sources(), 7)
    cpu = None
    if target:
        t = time.perf_counter()
        want = OracleGraph(el.n, el.src, el.tgt, None, False).betweenness(threads=0)
        ct = time.perf_counter() - t
        cpu = {"value": round(el.n * el.m / ct / 1e9, 4), "unit": "GTEPS",
               "cores": available_cores(), "kind": "port", "seconds": round(ct, 4),
               "sample": f"all {el.n} sources, all host threads, CPU: {cpu_model()}",
               "check": "bit-exact" if np.array_equal(res, want) else "differs"}
    out["er_gnm_2000_10000_all_sources"] = {
        "value": round(el.n * el.m / sec / 1e9, 3), "unit": "GTEPS", "seconds": round(sec, 5),
        "steps": 7, "dtype": "f64", "n": el.n, "m": el.m, "sources": el.n,
        "roofline": sweep_roofline("bfs_backward_kernel<double, false, false>", d,
                                   8 * d["reached_arcs"] + 32 * d["reached_nodes"]),
        "cpu_baseline": cpu}
    print(f"# config er2000: {sec * 1e3:.2f} ms", flush=True)

    # Barabasi-Albert n = 200,000, integer weights 1-16, every node a source
    el = gen.ba_weighted()
    g = Graph(el.n, el.src, el.tgt, el.weights, False)
    g.betweenness_sources(np.arange(64), weighted=True)
    (res, d), sec = med(lambda: g.betweenness_sources(weighted=True), 1)
    cpu = None
    if target:
        o = OracleGraph(el.n, el.src, el.tgt, el.weights, False)
        cpu = cpu_leg(lambda smp: o.betweenness_sources(smp, weighted=True, threads=0),
                      lambda smp: g.betweenness_sources(smp, weighted=True),
                      np.arange(el.n), el.m, el.n, d["reached_arcs"], target)
    out["ba_200000_m4_int_weights_1_16_all_sources"] = {
        "value": round(el.n * el.m / sec / 1e9, 3), "unit": "GTEPS", "seconds": round(sec, 3),
        "steps": 1, "dtype": "f64", "n": el.n, "m": el.m, "sources": el.n,
        "levels": d["levels"],
        "roofline": sweep_roofline("dial_backward_kernel<double, false, IntPath, false>", d,
                                   12 * d["reached_arcs"] + 32 * d["reached_nodes"]),
        "cpu_baseline": cpu}
    print(f"# config ba200k: {sec:.2f} s", flush=True)

    # 4,096 Watts-Strogatz graphs in one call
    els = gen.ws_batch(4096, 11)
    graphs = [(e.n, e.src, e.tgt) for e in els]
    work = sum(e.n * e.m for e in els)
    betweenness_batch(graphs)
    (res, d), sec = med(lambda: betweenness_batch(graphs), 2)
    parts = res
    cpu = None
    if target:
        step = 64
        sel = list(range(0, len(els), step))
        t = time.perf_counter()
        want = [OracleGraph(els[i].n, els[i].src, els[i].tgt, None, False).betweenness(threads=0)
                for i in sel]
        ct = time.perf_counter() - t
        differ = sum(not np.array_equal(parts[i], w) for i, w in zip(sel, want))
        cw = sum(els[i].n * els[i].m for i in sel)
        cpu = {"value": round(cw / ct / 1e9, 4), "unit": "GTEPS", "cores": available_cores(),
               "kind": "port", "seconds": round(ct, 3),
               "sample": f"{len(sel)} of 4096 graphs (every {step}th), one at a time, all "
                         f"host threads, CPU: {cpu_model()}",
               "check": "bit-exact" if not differ else f"{differ} graphs differ"}
    k1 = 128
    t = time.perf_counter()
    one = [Graph(e.n, e.src, e.tgt, None, False).betweenness() for e in els[:k1]]
    per_graph_s = (time.perf_counter() - t) / sum(e.n * e.m for e in els[:k1]) * work
    out["watts_strogatz_4096_graphs_all_sources"] = {
        "value": round(work / sec / 1e9, 3), "unit": "GTEPS", "seconds": round(sec, 3),
        "steps": 2, "dtype": "f64", "graphs_per_s": round(4096 / sec, 1),
        "one_graph_per_launch": {"seconds_extrapolated": round(per_graph_s, 3),
                                 "timed_graphs": k1, "speedup": round(per_graph_s / sec, 2),
                                 "bit_identical": all(np.array_equal(a, b)
                                                      for a, b in zip(one, parts[:k1]))},
        "plans": plan_legs(els, graphs, parts),
        "roofline": sweep_roofline("small_batch_kernel<double>", d,
                                   16 * d["reached_arcs"] + 60 * d["reached_nodes"]),
        "cpu_baseline": cpu}
    print(f"# config ws4096: {sec:.3f} s", flush=True)

    # undirected R-MAT scale 22, the 8,192 sources drawn from seed 7
    el = gen.rmat(22, 16, directed=False, seed=1)
    src = gen.sample_sources(el.n, 8192, 7)
    g = Graph(el.n, el.src, el.tgt, None, False)
    # warm-up: a free-order call (4,096 sources) builds the graph's device
    # state once, as the headline's warm-up step does (CSR, hub tables, the
    # landmark table that orders sources)
    g.betweenness_sources(src[:4096])
    (r64, d), sec = med(lambda: g.betweenness_sources(src), 1)
    (r32, d32), sec32 = med(lambda: g.betweenness_sources(src, precision=1), 1)
    nz = r64 != 0
    err = np.abs(r32[nz] - r64[nz]) / np.abs(r64[nz])
    cpu = None
    if target:
        o = OracleGraph(el.n, el.src, el.tgt, None, False)
        cpu = cpu_leg(lambda smp: o.betweenness_sources(smp, threads=0),
                      lambda smp: g.betweenness_sources(smp), src, el.m, len(src),
                      d["reached_arcs"], target)
        del o
    out["rmat22_undirected_8192_sampled_sources"] = {
        "value": round(len(src) * el.m / sec / 1e9, 3), "unit": "GTEPS",
        "seconds": round(sec, 3), "steps": 1, "dtype": "f64", "n": el.n, "m": el.m,
        "sources": len(src),
        "fp32_delta": {"seconds": round(sec32, 3),
                       "gteps": round(len(src) * el.m / sec32 / 1e9, 3),
                       "speedup_vs_fp64": round(sec / sec32, 3),
                       "max_rel_err_vs_fp64": float(err.max()) if nz.any() else 0.0,
                       "mean_rel_err_vs_fp64": float(err.mean()) if nz.any() else 0.0},
        "roofline": sweep_roofline("bfs_backward_kernel<double, false, false>", d,
                                   8 * d["reached_arcs"] +
                                   sweep_node_bytes(len(src)) * d["reached_nodes"]),
        "cpu_baseline": cpu}
    print(f"# config rmat22: {sec:.3f} s (FP32 dependencies {sec32:.3f} s)", flush=True)
    return out


def calibration_copy(nbytes: int = 1 << 30) -> None:
    """One device-to-device copy of nbytes through the HIP runtime: in a
    rocprofv3 --pmc run its FETCH_SIZE / WRITE_SIZE readings scale the
    counters of the other kernels to bytes."""
    import ctypes
    from knappnet_amd import _native
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    err = ctypes.create_string_buffer(256)
    if _native.device().netbc_calibration_copy(b, a, nbytes, err, 256):
        raise RuntimeError(err.value.decode())
    print(f"# calibration copy: {nbytes} bytes by netbc_copy_kernel", flush=True)
    # random 8-byte reads, one per 128-byte line of the source
    if _native.device().netbc_calibration_gather(b, a, nbytes, 128, err, 256):
        raise RuntimeError(err.value.decode())
    print(f"# calibration gather: {nbytes // 128} random 8-byte reads (one per 128-byte line) "
          f"by netbc_gather_kernel", flush=True)
    hip.hipFree(a)
    hip.hipFree(b)


def load_traffic(workload: str, precision: int):
    """Measured HBM bytes per launch of the dependency sweep, from the
    committed rocprofv3 --pmc summary for this workload (or None)."""
    path = ROOT / "profiles" / f"traffic_{workload}_p{precision}.json"
    if not path.exists():
        return None
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat18")
    ap.add_argument("--precision", type=int, default=0, help="0 FP64, 1 FP32 dependencies")
    ap.add_argument("--slots", type=int, default=0)
    ap.add_argument("--direction", type=int, default=0,
                    help="0 automatic, 1 top-down only, 2 bottom-up at every level")
    ap.add_argument("--alpha", type=float, default=0.0,
                    help="bottom-up when frontier out-arcs x alpha > unreached in-arcs (0: default)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a 1 GiB device copy first (scales rocprofv3 --pmc counters)")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the FP32-dependency run and the plain FP64 run")
    ap.add_argument("--max-sources", type=int, default=0,
                    help="time only the first K sources (experiments; not the headline)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the sampled baseline (0 = skip)")
    ap.add_argument("--edge", action="store_true",
                    help="edge betweenness (net_edge_betweenness) instead of node betweenness")
    ap.add_argument("--subset", action="store_true",
                    help="betweenness for sources and targets (every third node a target)")
    ap.add_argument("--closeness", action="store_true",
                    help="closeness of every node (net_closeness, normalized) instead")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the timed runs of the other workloads (headline only)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part "
                                   "of it (sources would shard across GPUs with one final "
                                   "all-reduce, see the design notes)", "n_gpus": args.gpus}))
        sys.exit(2)

    from knappnet_amd.engine import Graph, device_info

    if args.calibrate_copy:
        calibration_copy()
    if args.workload == "ws4096":
        run_ws_batch(args)
        return

    t0 = time.perf_counter()
    el, sources, config = build_workload(args.workload)
    gen_s = time.perf_counter() - t0
    weighted = el.weights is not None
    if args.max_sources:
        pool = np.arange(el.n) if sources is None else sources
        sources = np.asarray(pool[: args.max_sources])
        config = dict(config, workload=config["workload"] + f"_first{len(sources)}")
    if args.edge:
        config = dict(config, workload=config["workload"] + "_edges")
    if args.subset:
        config = dict(config, workload=config["workload"] + "_targets_every_third",
                      rescale="networkx_subset")
    if args.closeness:
        if sources is not None:
            raise SystemExit("closeness runs take every node as a source")
        config = dict(config, workload=config["workload"] + "_closeness", normalized=True,
                      unreachable="ignore")
    g = Graph(el.n, el.src, el.tgt, el.weights, el.directed)
    run = g.edge_betweenness_sources if args.edge else g.betweenness_sources
    targets = np.arange(0, el.n, 3)
    if args.subset:
        def run(sources=None, weighted=False, precision=0, slots=0, direction=0, alpha=0.0):
            nodes, _, d = g.betweenness_subset(
                sources=sources, targets=targets, weighted=weighted, rescale=2, edges=False,
                options=dict(precision=precision, slots=slots, direction=direction,
                             alpha=alpha))
            return nodes, d
    if args.closeness:
        def run(sources=None, weighted=False, precision=0, slots=0, direction=0, alpha=0.0):
            return g.closeness(weighted=weighted, normalized=True,
                               options=dict(slots=slots, direction=direction, alpha=alpha))
    k = el.n if sources is None else len(sources)
    print(f"# device: {device_info()}; workload {config['workload']} n={el.n} m={el.m} "
          f"sources={k} (generated in {gen_s:.1f}s)", flush=True)

    kw = dict(sources=sources, weighted=weighted, precision=args.precision, slots=args.slots,
              direction=args.direction, alpha=args.alpha)
    for i in range(args.warmup):
        t = time.perf_counter()
        res, diag = run(**kw)
        print(f"# warmup {i}: {time.perf_counter() - t:.3f}s build_ms={diag['build_ms']:.1f}",
              flush=True)

    times, diags = [], []
    for i in range(args.steps):
        t = time.perf_counter()
        res, diag = run(**kw)
        dt = time.perf_counter() - t
        times.append(dt)
        diags.append(diag)
        print(f"# step {i}: {dt:.3f}s kernel_ms={diag['kernel_ms']:.1f} "
              f"sweep_ms={diag['dominant_ms']:.1f} levels={diag['levels']} "
              f"bottom_up_steps={diag['bottom_up_levels']}", flush=True)
        if any(diag["work"]):
            print(f"# work counters: {diag['work']} reached_arcs={diag['reached_arcs']} "
                  f"reached_nodes={diag['reached_nodes']}", flush=True)
    total = sum(times)
    sec = total / len(times)
    gteps = k * el.m / sec / 1e9

    d = diags[-1]
    sweep_bytes = 8 * d["reached_arcs"] + sweep_node_bytes(k) * d["reached_nodes"]
    if args.closeness:
        sweep_bytes = d["dominant_bytes"]
    sweep_s = d["dominant_ms"] / 1e3
    achieved = sweep_bytes / sweep_s / 1e9 if sweep_s > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "kernel": "bfs_pull_kernel<false, false>" if args.closeness else
                  f"{'dial' if weighted else 'bfs'}_backward_kernel<"
                  f"{'double' if args.precision == 0 else 'float'}, "
                  f"{'true' if args.edge else 'false'}, "
                  f"{('RealPath, ' if args.workload == 'ba200k-real' else 'IntPath, ') if weighted else ''}"
                  f"false>",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": None,
        "traffic_source": None,
        "launches": d["dominant_launches"],
        "avg_launch_us": round(d["dominant_ms"] * 1e3 / max(d["dominant_launches"], 1), 2),
        "algorithmic_bytes_per_run": int(sweep_bytes),
    }

    tr = load_traffic(config["workload"], args.precision)
    if tr:
        roofline["traffic"] = tr.get("bytes_per_launch")
        roofline["traffic_source"] = tr.get("source")
        # the same launches priced by their measured HBM traffic
        if roofline["traffic"] and d["dominant_ms"] > 0:
            meas = roofline["traffic"] * d["dominant_launches"] / (d["dominant_ms"] / 1e3) / 1e9
            roofline["achieved_measured"] = round(meas, 2)
            roofline["frac_measured"] = round(meas / HBM_PEAK_GBS, 5)

    cpu = None
    if args.cpu_seconds > 0:
        from oracle.oracle import OracleGraph, available_cores
        o = OracleGraph(el.n, el.src, el.tgt, el.weights, el.directed)
        orun = o.edge_betweenness_sources if args.edge else o.betweenness_sources
        if args.subset:
            def orun(sample, weighted=False, threads=0):
                return o.betweenness_subset(sources=sample, targets=targets, weighted=weighted,
                                            rescale=2)[0]
        if args.closeness:
            def orun(sample, weighted=False, threads=0):
                p = o.distance_profiles(sample, weighted=weighted, threads=threads)
                with np.errstate(divide="ignore", invalid="ignore"):
                    return np.where(p["reached"] > 0, p["reached"] / p["total"], np.nan)
        pool = np.arange(el.n) if sources is None else np.asarray(sources)
        cores = 1 if args.subset else available_cores()
        take = min(len(pool), 256 if not args.subset else 16)
        while True:
            stride = max(1, len(pool) // take)
            sample = pool[::stride][:take]
            t = time.perf_counter()
            want = orun(sample, weighted=weighted, threads=0)
            ct = time.perf_counter() - t
            if ct >= args.cpu_seconds / 4 or take >= len(pool):
                break
            take = min(len(pool), int(take * max(2.0, args.cpu_seconds / 2 / max(ct, 1e-3))))
        # best of two timed runs of the final sample (the host may be shared)
        t = time.perf_counter()
        orun(sample, weighted=weighted, threads=0)
        ct = min(ct, time.perf_counter() - t)
        arcs_scale = None  # CPU time per reached arc, from the device's counts
        if args.closeness:
            got = np.nan_to_num(res[sample], nan=-1.0)
            want = np.nan_to_num(want, nan=-1.0)
        else:
            got, dsample = run(sample, weighted=weighted, precision=args.precision)
            if dsample["reached_arcs"] > 0:
                arcs_scale = d["reached_arcs"] / dsample["reached_arcs"]
        exact = bool(np.array_equal(got, want))
        nz = (got != 0) | (want != 0)
        rel = float(np.max(np.abs(got - want)[nz] / np.maximum(np.abs(want[nz]), 1e-300))) \
            if nz.any() else 0.0
        if arcs_scale:
            # the whole run's CPU time, scaled by reached arcs (sources differ in reach)
            cpu_gteps = k * el.m / (ct * arcs_scale) / 1e9
            how = "CPU time extrapolated to every source by reached arcs"
        else:
            cpu_gteps = len(sample) * el.m / ct / 1e9
            how = "rate per source"
        cpu = {"value": round(cpu_gteps, 4), "unit": "GTEPS", "cores": cores,
               "kind": "port", "seconds": round(ct, 3),
               "sample": f"{len(sample)} of {k} sources (every {stride}th), {cores} host "
                         f"thread{'s' if cores > 1 else ''}, {how}, CPU: {cpu_model()}",
               "check": "bit-exact" if exact else f"max rel err {rel:.3e}"}
        print(f"# cpu sample: {len(sample)} sources in {ct:.2f}s -> {cpu_gteps:.3f} GTEPS; "
              f"device vs oracle on sample: {cpu['check']}", flush=True)

    precision_check = None
    if (not args.no_compare and args.precision == 0 and not args.edge and not args.subset
            and not args.closeness):
        # The batched FP32-dependency form against FP64 (error) and against
        # the plain FP64 form: sources in the reference's order and sums in
        # its order, bit-identical to the CPU engine.
        (r32, d32), t32 = med(lambda: g.betweenness_sources(
            sources=sources, weighted=weighted, precision=1, slots=args.slots,
            direction=args.direction), 2)
        t = time.perf_counter()
        rplain, _ = g.betweenness_sources(sources=sources, weighted=weighted, precision=0,
                                          exact=True, slots=args.slots, direction=args.direction)
        tplain = time.perf_counter() - t
        nz = rplain != 0
        rel = np.abs(r32[nz] - rplain[nz]) / np.abs(rplain[nz])
        relb = np.abs(res[nz] - rplain[nz]) / np.abs(rplain[nz])
        precision_check = {
            "fp32_delta": {"seconds": round(t32, 3), "gteps": round(k * el.m / t32 / 1e9, 3),
                           "max_rel_err_vs_fp64": float(rel.max()) if nz.any() else 0.0,
                           "mean_rel_err_vs_fp64": float(rel.mean()) if nz.any() else 0.0},
            "plain_fp64": {"seconds": round(tplain, 3),
                           "gteps": round(k * el.m / tplain / 1e9, 3),
                           "form": "sources in the reference's order, sums in its order"},
            "batched_fp64_vs_plain_max_rel_err": float(relb.max()) if nz.any() else 0.0,
            "fp32_speedup_vs_plain_fp64": round(tplain / t32, 3),
            "fp32_speedup_vs_batched_fp64": round(sec / t32, 3),
        }
        # The target (FP32-delta form >= 1.5x the plain GPU form) read against
        # the FP64 form this line reports: the forward search keeps FP64 path
        # counts in either form, so only the sweep's share s of the FP64 run
        # can shrink; even a sweep that cost nothing would give 1 / (1 - s).
        share = d["dominant_ms"] / max(d["kernel_ms"], 1e-9)
        precision_check["north_star"] = {
            "ratio": round(sec / t32, 3), "against": "batched FP64 form (this line's value)",
            "target": 1.5, "met": sec / t32 >= 1.5,
            "fp64_sweep_share": round(share, 3),
            "ceiling_if_the_sweep_cost_nothing": round(1.0 / max(1.0 - share, 1e-9), 3),
            "fp32_sweep_ms": round(d32["dominant_ms"], 1), "fp64_sweep_ms": round(d["dominant_ms"], 1)}
        print(f"# precision check: {precision_check}", flush=True)

    line = {
        "metric": (METRIC_EDGE if args.edge else METRIC_SUBSET if args.subset
                   else METRIC_CLOSENESS if args.closeness else METRIC),
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f32-delta/f64-sigma",
        "data": "synthetic",
        "config": dict(config, n=el.n, m=el.m, sources=k, directed=el.directed,
                       weighted=weighted, slots=d["slots"]),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": round(gteps / cpu["value"], 1) if cpu else None,
        "precision_check": precision_check,
        "sigma_overflow": d["sigma_overflow"],
        "search": {"levels": d["levels"], "bottom_up_steps": d["bottom_up_levels"]},
    }
    if (args.workload == "rmat18" and not (args.edge or args.subset or args.closeness)
            and args.precision == 0 and not args.max_sources and not args.no_configs):
        line["configs"] = extra_configs(args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
