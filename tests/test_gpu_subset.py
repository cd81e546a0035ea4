"""Betweenness for sources and targets on the device (net_betweenness_subset),
against the CPU oracle's restatement of path_centrality.rs:421-689.

The bar: within 1e-12 relative of the oracle (the reference adds each
source's dependency terms per predecessor arc in settle order, the device
per node in row order), and within the reference's tolerance (rel 1e-9,
abs 1e-12) of the committed definition and NetworkX goldens, including
edge values with NetworkX's subset edge split.
"""
import sys

import numpy as np
import pytest

from _common import (SUBSET_VARIANTS, close, golden_values, random_case, subset_call,
                     subset_networkx_documented)
from knappnet_amd import generate as gen
from knappnet_amd.engine import EngineError, Graph
from oracle.oracle import OracleGraph

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_rel(a, b, what, tol=1e-12, abs_=1e-300):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, what
    assert np.array_equal(np.isnan(a), np.isnan(b)), f"{what}: NaN pattern differs"
    keep = ~np.isnan(a)
    a, b = a[keep], b[keep]
    err = np.abs(a - b) / np.maximum(np.abs(b), abs_)
    err[np.abs(a - b) <= abs_] = 0.0
    assert float(err.max(initial=0.0)) <= tol, f"{what}: max rel err {err.max():.3e}"


def both(n, s, t, w, directed):
    return Graph(n, s, t, w, directed), OracleGraph(n, s, t, w, directed)


def test_subset_goldens(fixtures):
    for variant, spec in SUBSET_VARIANTS.items():
        for name, fx in fixtures.items():
            if variant not in fx["golden"]:
                continue
            kw, out = subset_call(fx["n"], spec)
            say("fixture", name, variant)
            g, o = both(fx["n"], fx["src"], fx["tgt"], fx["weights"], fx["directed"])
            nodes, edges = g.betweenness_subset(**kw)
            onodes, oedges = o.betweenness_subset(**kw)
            got, want = (nodes, onodes) if out == "nodes" else (edges, oedges)
            assert_rel(got, want, f"{name} {variant} vs oracle", abs_=1e-13)
            for oracle, gold in fx["golden"][variant].items():
                if oracle == "networkx" and subset_networkx_documented(fx, variant):
                    continue
                assert close(got, golden_values(gold), 1e-9, 1e-12), (name, variant, oracle)


@pytest.mark.parametrize("seed", range(8))
def test_subset_random(seed):
    rng = np.random.default_rng(700 + seed)
    for case in range(5):
        n, s, t, directed = random_case(rng, max_n=70)
        w_int = rng.integers(1, 5, len(s)).astype(np.float64)
        w_real = rng.random(len(s)) * 3 + 0.5
        srcs = list(rng.integers(0, n, size=int(rng.integers(1, 2 * n + 1))))
        tg = list(rng.choice(n, size=int(rng.integers(1, n + 1)), replace=False))
        say("seed", seed, "case", case, n, len(s), directed)
        for w, weighted in ((None, False), (w_int, True), (w_real, True)):
            g, o = both(n, s, t, w, directed)
            for kw in (dict(sources=srcs, targets=tg),
                       dict(sources=srcs, targets=tg, endpoints=True, normalized=True),
                       dict(targets=tg, parallel=1, tie=1, rescale=2),
                       dict(sources=srcs[:5], parallel=1, tie=1, rescale=1, normalized=True),
                       dict(sources=srcs[:5], parallel=1, rescale=1, endpoints=True),
                       dict(sources=srcs, tie=1, rescale=0, normalized=True, endpoints=True),
                       dict(sources=srcs, targets=tg, edge_split=1),
                       dict(sources=srcs[:9], targets=tg, edge_split=1, parallel=1, tie=1,
                            rescale=2)):
                kw = dict(kw, weighted=weighted)
                nodes, edges = g.betweenness_subset(**kw)
                onodes, oedges = o.betweenness_subset(**kw)
                assert_rel(nodes, onodes, f"nodes {kw}", abs_=1e-13)
                assert_rel(edges, oedges, f"edges {kw}", abs_=1e-13)


def test_subset_every_pair_is_betweenness():
    el = gen.rmat(12, 8, directed=True, seed=2)
    g = Graph(el.n, el.src, el.tgt, None, True)
    nodes, edges = g.betweenness_subset()
    assert_rel(nodes, g.betweenness(), "nodes", tol=1e-12)
    assert_rel(edges, g.edge_betweenness(), "edges", tol=1e-12)


def test_subset_large_sampled():
    # 4,096 or more sources (free order), targets, endpoints, on R-MAT 13
    el = gen.rmat(13, 8, directed=False, seed=5)
    g, o = both(el.n, el.src, el.tgt, None, False)
    rng = np.random.default_rng(3)
    srcs = rng.integers(0, el.n, 5000)
    tg = rng.choice(el.n, el.n // 3, replace=False)
    say("rmat13 sampled", el.n, el.m)
    for kw in (dict(sources=srcs, targets=tg, endpoints=True),
               dict(sources=srcs, rescale=1, normalized=True, parallel=1, tie=1),
               dict(sources=srcs, targets=tg, edge_split=1)):
        nodes, edges = g.betweenness_subset(**kw)
        onodes, oedges = o.betweenness_subset(**kw)
        assert_rel(nodes, onodes, f"nodes {kw}")
        assert_rel(edges, oedges, f"edges {kw}")


def test_subset_errors():
    g = Graph(3, [0, 1], [1, 2], [1.0, -2.0], True)
    with pytest.raises(EngineError, match=r"^node index 5 is out of range for a network "
                                          r"with 3 nodes$"):
        g.betweenness_subset(targets=[0, 5])
    for kw, msg in (({"rescale": 3}, "unknown rescaling code 3"),
                    ({"parallel": 2}, "unknown code 2 for parallel edges"),
                    ({"edge_split": -1}, "unknown edge split code -1"),
                    ({"tie": 2}, "unknown tie code 2"),
                    ({"weights_as": 4}, "unknown weight kind code 4")):
        with pytest.raises(EngineError, match="^" + msg + "$"):
            g.betweenness_subset(**kw)
    with pytest.raises(EngineError, match=r"^edge 1 has length -2; shortest-path betweenness "
                                          r"needs strictly positive lengths$"):
        g.betweenness_subset(weighted=True)
    with pytest.raises(EngineError, match=r"^edge 1 has weight -2; weights read as strengths "
                                          r"become lengths 1/w and must be strictly positive$"):
        g.betweenness_subset(weighted=True, weights_as="strength")
    with pytest.raises(EngineError, match=r"^the graph has no edge weights$"):
        Graph(2, [0], [1], None, True).betweenness_subset(weighted=True)
    # node values do not depend on the edge split
    nodes, edges = Graph(3, [0, 1], [1, 2], None, True).betweenness_subset(edge_split=1,
                                                                          targets=[2])
    assert list(nodes) == [0.0, 1.0, 0.0] and list(edges) == [1.0, 2.0]
    w = [1e-4, 1.0, 10.0]
    with pytest.raises(EngineError, match=r"exceeds 2,047 times the shortest is not available"):
        Graph(4, [0, 1, 2], [1, 2, 3], w, True).betweenness_subset(weighted=True,
                                                                  edge_split=1)
