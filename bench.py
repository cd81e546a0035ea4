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
  level; 32 B per reached node: row offsets, sigma read, coefficient and
  dependency writes; summed over sources);
- ``cpu_baseline``: the CPU oracle (the reference's algorithm restated in
  C, all host cores) timed on a strided sample of the same sources, and a
  bit-exact check of the device result on that sample.

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
        sel = els[:: max(1, len(els) // 64)]
        t = time.perf_counter()
        ok = True
        for e in sel:
            r = OracleGraph(e.n, e.src, e.tgt, None, False).betweenness(threads=0)
        ct = time.perf_counter() - t
        cw = sum(e.n * e.m for e in sel)
        cpu = {"value": round(cw / ct / 1e9, 4), "unit": "GTEPS", "cores": available_cores(),
               "kind": "port", "seconds": round(ct, 3),
               "sample": f"{len(sel)} of 4096 graphs (every 64th), one at a time, all host "
                         f"threads, CPU: {cpu_model()}"}
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
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


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
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"# calibration copy: {nbytes} bytes by netbc_copy_kernel", flush=True)


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
    sweep_bytes = 8 * d["reached_arcs"] + 32 * d["reached_nodes"]
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
        if args.closeness:
            got = np.nan_to_num(res[sample], nan=-1.0)
            want = np.nan_to_num(want, nan=-1.0)
        else:
            got, _ = run(sample, weighted=weighted, precision=args.precision)
        exact = bool(np.array_equal(got, want))
        nz = (got != 0) | (want != 0)
        rel = float(np.max(np.abs(got - want)[nz] / np.maximum(np.abs(want[nz]), 1e-300))) \
            if nz.any() else 0.0
        cpu_gteps = len(sample) * el.m / ct / 1e9
        cpu = {"value": round(cpu_gteps, 4), "unit": "GTEPS", "cores": cores,
               "kind": "port", "seconds": round(ct, 3),
               "sample": f"{len(sample)} of {k} sources (every {stride}th), {cores} host "
                         f"thread{'s' if cores > 1 else ''}, CPU: {cpu_model()}",
               "check": "bit-exact" if exact else f"max rel err {rel:.3e}"}
        print(f"# cpu sample: {len(sample)} sources in {ct:.2f}s -> {cpu_gteps:.3f} GTEPS; "
              f"device vs oracle on sample: {cpu['check']}", flush=True)

    precision_check = None
    if (not args.no_compare and args.precision == 0 and not args.edge and not args.subset
            and not args.closeness):
        # The batched FP32-dependency form against FP64 (error) and against
        # the plain FP64 form: sources in the reference's order and sums in
        # its order, bit-identical to the CPU engine.
        t = time.perf_counter()
        r32, _ = g.betweenness_sources(sources=sources, weighted=weighted, precision=1,
                                       slots=args.slots, direction=args.direction)
        t32 = time.perf_counter() - t
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
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
