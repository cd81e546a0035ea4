"""Real path lengths on the device (betweenness.rs:297-368), through the C ABI.

Two engines: the bucketed search (buckets of half the shortest arc, 64
sources per batch; used when the longest arc is at most about 2,047 times
the shortest) and, for more unequal lengths, the reference's Dijkstra with
its 4-ary heap, one source per thread. The bar: bit-exact against the
oracle where no two path lengths tie within 1e-10 without being equal
(random real weights), and within 1e-12 relative where floating-point sums
of equal real lengths differ in their last bits (the reference keeps the
first such length its heap meets, the bucketed search the least).
"""
import sys

import numpy as np
import pytest

from _common import random_case
from knappnet_amd import generate as gen
from knappnet_amd.engine import Graph
from oracle.oracle import OracleGraph

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(n, s, t, w, directed):
    return Graph(n, s, t, w, directed), OracleGraph(n, s, t, w, directed)


def assert_same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    if not np.array_equal(a, b):
        i = int(np.argmax(a != b))
        raise AssertionError(f"{what}: first difference at {i}: {a[i]!r} vs {b[i]!r}; "
                             f"{int(np.sum(a != b))} differ")


def assert_rel(a, b, what, tol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    err = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    err[(a == 0) & (b == 0)] = 0.0
    assert float(err.max(initial=0.0)) <= tol, f"{what}: max rel err {err.max():.3e}"


@pytest.mark.parametrize("seed", range(6))
def test_bucketed_random_multigraphs(seed):
    rng = np.random.default_rng(40 + seed)
    for case in range(6):
        n, s, t, directed = random_case(rng, max_n=120)
        w = rng.random(len(s)) * 7.5 + 0.5
        say("seed", seed, "case", case, n, len(s), directed)
        g, o = pair(n, s, t, w, directed)
        assert_same(g.betweenness(weighted=True), o.betweenness(weighted=True), "distance")
        assert_same(g.betweenness(weighted=True, weights_as="strength"),
                    o.betweenness(weighted=True, weights_as="strength"), "strength")
        assert_same(g.betweenness(weighted=True, normalized=True),
                    o.betweenness(weighted=True, normalized=True), "normalized")
        assert_rel(g.edge_betweenness(weighted=True), o.edge_betweenness(weighted=True),
                   "edge")


def test_near_ties():
    # Lengths whose sums tie in exact arithmetic but not in floating point
    # (0.1 + 0.2 != 0.3): the reference counts them as ties (relative 1e-10).
    rng = np.random.default_rng(7)
    choices = np.array([0.1, 0.2, 0.3, 0.05, 0.15, 0.25])
    for case in range(10):
        n, s, t, directed = random_case(rng, max_n=60)
        w = rng.choice(choices, size=len(s))
        say("ties", case, n, len(s), directed)
        g, o = pair(n, s, t, w, directed)
        assert_rel(g.betweenness(weighted=True), o.betweenness(weighted=True), "ties")
        assert_rel(g.edge_betweenness(weighted=True), o.edge_betweenness(weighted=True),
                   "edge ties")
    g, o = pair(4, [0, 1, 0, 2], [1, 3, 2, 3], [0.1, 0.2, 0.15, 0.15], True)
    r = g.betweenness(weighted=True)
    assert abs(r[1] - 0.5) < 1e-12 and abs(r[2] - 0.5) < 1e-12


def test_bucketed_ba_exact_and_free_order():
    el = gen.preferential(3000, 4, 5, 0)
    w = np.random.default_rng(1).random(el.m) * 3.0 + 1.0
    g, o = pair(el.n, el.src, el.tgt, w, False)
    say("ba3000 real all sources")
    got, diag = g.betweenness_sources(weighted=True)
    assert diag["path_kind"] == 2
    assert_same(got, o.betweenness(weighted=True, threads=0), "ba3000")
    # 4,096 or more sources: batched by landmark distances (free order)
    src = np.concatenate([np.arange(el.n), np.arange(el.n)[::2]])  # 4,500 sources
    got, _ = g.betweenness_sources(src, weighted=True)
    assert_rel(got, o.betweenness_sources(src, weighted=True, threads=0), "free order")
    f32, _ = g.betweenness_sources(weighted=True, precision=1)
    assert_rel(f32, o.betweenness(weighted=True, threads=0), "fp32", tol=1e-6)


def test_very_unequal_lengths_use_the_per_thread_search():
    rng = np.random.default_rng(9)
    for case in range(4):
        n, s, t, directed = random_case(rng, max_n=50)
        w = 10.0 ** rng.uniform(-4, 1, len(s))  # longest / shortest up to 1e5
        say("unequal", case, n, len(s), directed)
        g, o = pair(n, s, t, w, directed)
        assert_same(g.betweenness(weighted=True), o.betweenness(weighted=True), "unequal")
        assert_rel(g.edge_betweenness(weighted=True), o.edge_betweenness(weighted=True),
                   "unequal edge")


def test_deep_real_paths():
    # a long path with unit-ish lengths: many buckets, most of them empty
    n = 400
    s = np.arange(n - 1)
    t = s + 1
    w = 1.0 + (np.arange(n - 1) % 7) * 0.5
    g, o = pair(n, s, t, w, True)
    assert_same(g.betweenness(weighted=True), o.betweenness(weighted=True), "path")
