"""Edge betweenness on the device (net_edge_betweenness and
netbc_edge_betweenness_sources) against the CPU oracle, through the C ABI.

The bar: within 1e-12 relative of the oracle's restatement of
reach.rs:174-232 (the device adds each arc's credit over its sources
before adding the arcs of an edge, the reference adds per source, so the
association differs); the reference's own tolerance against the committed
definition goldens (rel 1e-9, abs 1e-9 / 1e-12 normalised); results that do
not depend on the run (bit-identical repeats); FP32 dependencies within
1e-6 relative of FP64.
"""
import sys

import numpy as np
import pytest

from _common import EDGE_VARIANTS, close, random_case, tolerance
from knappnet_amd import generate as gen
from knappnet_amd.engine import EngineError, Graph
from oracle.oracle import OracleGraph

pytestmark = pytest.mark.gpu

REL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_rel(a, b, what, tol=REL):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    nan = np.isnan(a) | np.isnan(b)
    assert np.array_equal(np.isnan(a), np.isnan(b)), f"{what}: NaN pattern differs"
    a, b = a[~nan], b[~nan]
    err = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    err[(a == 0) & (b == 0)] = 0.0
    assert float(err.max(initial=0.0)) <= tol, f"{what}: max rel err {err.max():.3e}"


def test_edge_fixtures(fixtures):
    for name, fx in fixtures.items():
        g = Graph(fx["n"], fx["src"], fx["tgt"], fx["weights"], fx["directed"])
        o = OracleGraph(fx["n"], fx["src"], fx["tgt"], fx["weights"], fx["directed"])
        for variant, spec in EDGE_VARIANTS.items():
            if variant not in fx["golden"]:
                continue
            say("fixture", name, variant)
            got = g.edge_betweenness(**spec)
            assert_rel(got, o.edge_betweenness(**spec), f"{name} {variant}")
            rel, ab = tolerance(variant)
            assert close(got, fx["golden"][variant]["definition"], rel, ab), (name, variant)


def test_edge_known_answers():
    # reach.rs:731-826
    def eb(n, s, t, w=None, directed=False, **kw):
        return Graph(n, s, t, w, directed).edge_betweenness(**kw)

    assert list(eb(3, [0, 1], [1, 2])) == [2.0, 2.0]
    assert list(eb(3, [0, 1], [1, 2], directed=True)) == [2.0, 2.0]
    assert list(eb(3, [0, 0, 1, 2], [1, 1, 2, 2])) == [1.0, 1.0, 2.0, 0.0]
    sq = ([0, 1, 2, 3], [1, 2, 3, 0])
    assert list(eb(4, *sq)) == [2.0] * 4
    assert close(eb(4, *sq, normalized=True), [2.0 / 6.0] * 4, 1e-15, 1e-15)
    assert np.isnan(eb(1, [0], [0], normalized=True)).all()
    assert eb(0, [], []).size == 0
    assert eb(5, [], []).size == 0
    e = [(0, 1), (0, 2), (1, 3), (2, 3), (1, 4), (3, 5), (4, 5)]
    w = [2.0, 1.0, 1.0, 2.0, 3.0, 2.0, 1.0]
    s, t = zip(*e)
    g = Graph(6, s, t, w, True)
    o = OracleGraph(6, s, t, w, True)
    assert_rel(g.edge_betweenness(weighted=True), o.edge_betweenness(weighted=True), "dial")
    assert_rel(g.edge_betweenness(weighted=True, weights_as="strength"),
               o.edge_betweenness(weighted=True, weights_as="strength"), "dijkstra")


def test_edge_error_paths():
    g = Graph(2, [0], [1], [0.0], True)
    with pytest.raises(EngineError, match=r"^edge 0 has weight 0; this algorithm treats "
                                          r"weights as distances and requires them to be "
                                          r"strictly positive$"):
        g.edge_betweenness(weighted=True)
    with pytest.raises(EngineError, match=r"^the graph has no edge weights$"):
        Graph(2, [0], [1], None, True).edge_betweenness(weighted=True)
    with pytest.raises(EngineError, match=r"^unknown weight kind code 7$"):
        Graph(2, [0], [1], None, True).edge_betweenness(weights_as=7)
    with pytest.raises(EngineError, match=r"^source 1 is node index 5"):
        Graph(3, [0], [1], None, True).edge_betweenness_sources([0, 5])


@pytest.mark.parametrize("seed", range(12))
def test_edge_random_multigraphs(seed):
    rng = np.random.default_rng(500 + seed)
    for case in range(8):
        n, s, t, directed = random_case(rng, max_n=60)
        w_int = rng.integers(1, 6, len(s)).astype(np.float64)
        w_real = rng.random(len(s)) + 0.25
        say("seed", seed, "case", case, n, len(s), directed)
        for w, kw in ((None, {}), (w_int, {"weighted": True}),
                      (w_real, {"weighted": True}),
                      (w_int, {"weighted": True, "weights_as": "strength"}),
                      (None, {"normalized": True})):
            g = Graph(n, s, t, w, directed)
            o = OracleGraph(n, s, t, w, directed)
            assert_rel(g.edge_betweenness(**kw), o.edge_betweenness(**kw), f"{kw}")


@pytest.mark.parametrize("directed", [True, False])
def test_edge_large_free_order(directed):
    # >= 4096 sources: batched by landmark distances, heavy rows split.
    el = gen.rmat(13, 16, directed=directed, seed=3)
    g = Graph(el.n, el.src, el.tgt, None, directed)
    o = OracleGraph(el.n, el.src, el.tgt, None, directed)
    say("rmat13", directed, el.n, el.m)
    got = g.edge_betweenness()
    want = o.edge_betweenness(threads=0)
    assert_rel(got, want, "rmat13 all sources")
    # deterministic: a second run is bit-identical
    assert np.array_equal(got, g.edge_betweenness())
    # the exact-order option changes only the node sums' order
    ex, _ = g.edge_betweenness_sources(exact=True)
    assert_rel(ex, want, "rmat13 exact order")
    f32, diag = g.edge_betweenness_sources(precision=1)
    assert_rel(f32, want, "rmat13 fp32", tol=1e-6)
    say("levels", diag["levels"], "slots", diag["slots"])


def test_edge_weighted_large():
    el = gen.ba_weighted(n=3000, m=4, seed=5)
    g = Graph(el.n, el.src, el.tgt, el.weights, el.directed)
    o = OracleGraph(el.n, el.src, el.tgt, el.weights, el.directed)
    say("ba", el.n, el.m)
    assert_rel(g.edge_betweenness(weighted=True), o.edge_betweenness(weighted=True, threads=0),
               "dial all sources")
    w = np.random.default_rng(1).random(el.m) + 0.5
    g = Graph(el.n, el.src, el.tgt, w, el.directed)
    o = OracleGraph(el.n, el.src, el.tgt, w, el.directed)
    src = np.arange(0, el.n, 7)
    got, _ = g.edge_betweenness_sources(src, weighted=True)
    assert_rel(got, o.edge_betweenness_sources(src, weighted=True, threads=0),
               "dijkstra sampled")


def test_edge_sources_with_repeats():
    el = gen.rmat(11, 8, directed=True, seed=9)
    g = Graph(el.n, el.src, el.tgt, None, True)
    o = OracleGraph(el.n, el.src, el.tgt, None, True)
    src = np.random.default_rng(2).integers(0, el.n, 5000)
    got, _ = g.edge_betweenness_sources(src)
    assert_rel(got, o.edge_betweenness_sources(src, threads=0), "sources with repeats")
    got, _ = g.edge_betweenness_sources(src[:100], normalized=True)
    assert_rel(got, o.edge_betweenness_sources(src[:100], normalized=True), "short list")


@pytest.mark.parametrize("directed", [True, False])
def test_edge_flow_identity_full_size(directed):
    # Every shortest path of length d crosses d edges and d - 1 interior
    # nodes, so sum(edge scores) - sum(node scores) counts the reachable
    # ordered pairs (halved undirected): reached nodes minus the sources.
    el = gen.rmat(16, 16, directed=directed, seed=4)
    g = Graph(el.n, el.src, el.tgt, None, directed)
    say("rmat16", directed, el.n, el.m)
    eb, d1 = g.edge_betweenness_sources()
    bc, d2 = g.betweenness_sources()
    pairs = d2["reached_nodes"] - el.n
    assert d1["reached_nodes"] == d2["reached_nodes"]
    want = pairs if directed else pairs / 2
    assert abs((eb.sum() - bc.sum()) - want) <= 1e-9 * want
    # per node (directed): what flows in is b(v) plus the sources reaching v
    if directed:
        inflow = np.zeros(el.n)
        np.add.at(inflow, el.tgt, eb)
        reach_in = inflow - bc
        assert np.all(np.abs(reach_in - np.round(reach_in)) <= 1e-6 * np.maximum(reach_in, 1))
        assert abs(np.round(reach_in).sum() - pairs) <= 0
