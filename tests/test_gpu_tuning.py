"""The engine settings an autotuner chooses, through the C ABI.

Every setting changes only how the work is cut (workgroup sizes, which
nodes count as hubs and how their in-rows are split, how many integer
distances share a bucket), never the arithmetic: in exact order the scores
are the oracle's bit for bit under any setting, and in the default order
the levels and path counts stay bit-exact. Out-of-range settings are
refused with a message. Reference arithmetic: crates/engine/src/metrics/
betweenness.rs:161-295 (searches and sweep), parallel.rs:383-503 (the
ordered cross-source sum).
"""
import sys

import numpy as np
import pytest

from knappnet_amd import generate as gen
from knappnet_amd.engine import EngineError, Graph, betweenness_batch
from oracle.oracle import OracleGraph

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_same(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    if not np.array_equal(a, b):
        i = int(np.argmax(a != b))
        raise AssertionError(f"{what}: first difference at {i}: {a[i]!r} vs {b[i]!r}; "
                             f"{int(np.sum(a != b))} differ")


def test_settings_are_checked():
    g = Graph(4, [0, 1, 2], [1, 2, 3], None, False)
    bad = [
        (dict(sweep_block=100), r"^sweep block 100 is not 0 or a multiple of 64 up to 1024$"),
        (dict(step_block=2048), r"^step block 2048 is not 0 or a multiple of 64 up to 1024$"),
        (dict(hub_in=-1), r"^hub in-degree -1 is outside 0..2\^30$"),
        (dict(hub_chunk=100), r"^hub chunk 100 is not 0 or a multiple of 64 up to 65536$"),
        (dict(bucket_width=-2), r"^negative bucket width$"),
    ]
    for kw, msg in bad:
        with pytest.raises(EngineError, match=msg):
            g.betweenness_sources(**kw)
    # the handle keeps working
    assert list(g.betweenness_sources()[0]) == [0.0, 2.0, 2.0, 0.0]


@pytest.mark.parametrize("directed", [True, False])
def test_block_sizes_and_hub_cuts_keep_the_bits(directed):
    """R-MAT scale 12 (hubs and heavy rows) in every search direction with
    sweep and step workgroups of 64 to 1024 threads and hub thresholds and
    chunk sizes from very small (most nodes are hubs, rows cut into 64-arc
    tasks) to none: exact order is the oracle's bits; changing the hub set
    between runs on one handle starts the new hubs' rows from zero."""
    el = gen.rmat(12, 32, directed=directed, seed=9)
    g = Graph(el.n, el.src, el.tgt, None, directed)
    want = OracleGraph(el.n, el.src, el.tgt, None, directed).betweenness(threads=0)
    settings = [
        dict(sweep_block=64, step_block=1024),
        dict(sweep_block=1024, step_block=64),
        dict(sweep_block=512, step_block=512, hub_in=1, hub_chunk=64),
        dict(hub_in=4, hub_chunk=128, direction=2),
        dict(hub_in=1 << 30, hub_chunk=65536, direction=2),
        dict(hub_in=16, hub_chunk=1024, direction=1),
        dict(),
    ]
    for kw in settings:
        say("settings", directed, kw)
        got, _ = g.betweenness_sources(exact=True, **kw)
        assert_same(got, want, f"exact order under {kw}")
    # default order: deterministic for a setting, within 1e-12 of the oracle
    a, _ = g.betweenness_sources(np.arange(el.n).repeat(2), sweep_block=128, hub_in=8)
    b, _ = g.betweenness_sources(np.arange(el.n).repeat(2), sweep_block=128, hub_in=8)
    assert_same(a, b, "free order is deterministic")
    rel = np.abs(a - 2 * want) / np.maximum(np.abs(2 * want), 1e-300)
    rel[(a == 0) & (want == 0)] = 0.0
    assert rel.max() <= 1e-12


def test_bucket_width_on_integer_lengths():
    """Integer lengths 4-12: buckets of 1, 2 and 4 distances (up to the
    shortest arc) give the oracle's distances, path counts and scores bit
    for bit; a wider request is cut to the shortest arc; with lengths from
    1 the width stays 1."""
    el = gen.ba_weighted(n=3000, m=4, seed=5)
    w = 4.0 + np.floor(el.weights * 8.0 / 16.0)  # integers 4..12
    g = Graph(el.n, el.src, el.tgt, w, False)
    o = OracleGraph(el.n, el.src, el.tgt, w, False)
    want = o.betweenness(weighted=True, threads=0)
    levels = {}
    for width in (1, 2, 4, 9):
        got, diag = g.betweenness_sources(weighted=True, exact=True, bucket_width=width)
        say("bucket width", width, "levels", diag["levels"])
        assert_same(got, want, f"bucket width {width}")
        levels[width] = diag["levels"]
    assert levels[2] < levels[1] and levels[4] < levels[2]
    assert levels[9] == levels[4]  # cut to the shortest arc (4)
    srcs = np.array([0, 1, 17, 2999])
    dist, sig, _ = g.state_dump(srcs, weighted=True)
    for i, s in enumerate(srcs):
        _, du, _, osig, _, _ = o.single_source(s, weighted=True)
        assert_same(dist[i], du, f"distances from {s}")
        assert_same(sig[i], osig, f"path counts from {s}")
    # free order (landmark batches) with wide buckets: within 1e-12
    many = np.arange(el.n).repeat(2)
    got, _ = g.betweenness_sources(many, weighted=True, bucket_width=4)
    ref = 2 * want
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    rel[(got == 0) & (ref == 0)] = 0.0
    assert rel.max() <= 1e-12
    # lengths from 1: any requested width runs as 1
    g1 = Graph(el.n, el.src, el.tgt, el.weights, False)
    a, d1 = g1.betweenness_sources(weighted=True, exact=True, bucket_width=8)
    b, d2 = g1.betweenness_sources(weighted=True, exact=True, bucket_width=1)
    assert_same(a, b, "width cut to 1")
    assert d1["levels"] == d2["levels"]


def test_many_graphs_workgroup_sizes():
    """The one-workgroup engine at 256, 512 and 1024 threads per workgroup:
    the same bits (each graph's scores equal one net_betweenness call per
    graph); other sizes are refused."""
    els = gen.ws_batch(48, 21)
    graphs = [(e.n, e.src, e.tgt) for e in els]
    want, _ = betweenness_batch(graphs, sweep_block=512)
    for block in (256, 1024, 0):
        got, _ = betweenness_batch(graphs, sweep_block=block)
        for i, (a, b) in enumerate(zip(got, want)):
            assert_same(a, b, f"graph {i}, {block} threads")
    for i in (0, 17, 47):
        e = els[i]
        one = Graph(e.n, e.src, e.tgt, None, False).betweenness()
        assert_same(want[i], one, f"graph {i} against one call")
    with pytest.raises(EngineError, match=r"^the one-workgroup engine runs workgroups of 256, "
                                          r"512 or 1024 threads, not 128$"):
        betweenness_batch(graphs, sweep_block=128)


def test_slots_even_out_over_super_batches():
    """65 batches of 64 sources with 64 slots requested run as 2 super-
    batches of 33 slots (not 64 and 1): the diagnostics report 33, exact
    order is the oracle's bits, free order within 1e-12."""
    el = gen.rmat(12, 16, directed=True, seed=4)
    g = Graph(el.n, el.src, el.tgt, None, True)
    src = np.arange(65 * 64) % el.n
    want = OracleGraph(el.n, el.src, el.tgt, None, True).betweenness_sources(src, threads=0)
    got, diag = g.betweenness_sources(src, exact=True, slots=64)
    assert diag["slots"] == 33
    assert_same(got, want, "exact order, 33 slots")
    free, diag = g.betweenness_sources(src, slots=64)
    assert diag["slots"] == 33
    rel = np.abs(free - want) / np.maximum(np.abs(want), 1e-300)
    rel[(free == 0) & (want == 0)] = 0.0
    assert rel.max() <= 1e-12
    # a request within one super-batch is kept
    _, diag = g.betweenness_sources(src[:40 * 64], slots=64)
    assert diag["slots"] == 40
