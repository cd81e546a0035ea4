"""Batch plans (netbc_batch_plan_*): one batch shape, replayed.

A plan fixes the graph count, node counts and edge counts; each run takes
new edges of that shape (the replicate batches of a conditional uniform
graph test, r/R/cug.R:25-47). With capture the device side (the union
CSR built from the edges, the persistent launch, the ordered reduction)
is one recorded HIP graph. Every replay is compared bit for bit with
netbc_betweenness_batch on the same graphs, with other device work run
between replays, and a sample of graphs with the CPU oracle.
"""
import sys

import numpy as np
import pytest

from knappnet_amd import generate as gen
from knappnet_amd.engine import BatchPlan, EngineError, Graph, betweenness_batch
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


def draw(rng, ns, ms, deep):
    """New edges of the shape (ns, ms): random multigraphs with loops and
    parallel edges, and for index `deep` a path in a random node order
    (a search past 254 levels, which the persistent launch hands back)."""
    graphs = []
    for i, (n, m) in enumerate(zip(ns, ms)):
        if i == deep:
            p = rng.permutation(n)
            graphs.append((n, p[:-1], p[1:]))
        else:
            graphs.append((n, rng.integers(0, max(n, 1), m), rng.integers(0, max(n, 1), m)))
    return graphs


@pytest.mark.parametrize("capture", [True, False])
@pytest.mark.parametrize("directed", [False, True])
def test_plan_replays_match_batch_entry(directed, capture):
    rng = np.random.default_rng(71 + 2 * directed + capture)
    ns = [int(rng.integers(1, 300)) for _ in range(30)] + [1, 2, 0, 2048, 2500, 600]
    ms = [int(rng.integers(0, 4 * n)) if n else 0 for n in ns]
    deep = len(ns) - 1
    ms[deep] = ns[deep] - 1
    plan = BatchPlan(ns, ms, directed=directed, capture=capture)
    other = Graph(50, np.arange(49), np.arange(1, 50), None, directed)
    last = None
    for rep in range(3):
        graphs = draw(rng, ns, ms, deep)
        say("plan", directed, capture, "replay", rep)
        got, diag = plan.run(graphs)
        want, _ = betweenness_batch(graphs, directed=directed)
        for i in range(len(ns)):
            assert_same(got[i], want[i], f"replay {rep} graph {i} (n={ns[i]})")
        assert diag["reached_nodes"] > 0
        other.betweenness()  # other device work between replays
        last = graphs
    for i in (0, 7, 30, 33, deep):
        n, s, t = last[i]
        want = OracleGraph(n, s, t, None, directed).betweenness(threads=0)
        assert_same(plan.run(last)[0][i], want, f"oracle graph {i}")
    plan.close()


def test_plan_normalized_and_fp32_dependencies():
    rng = np.random.default_rng(5)
    ns = [int(rng.integers(2, 400)) for _ in range(24)]
    ms = [3 * n for n in ns]
    for precision in (0, 1):
        plan = BatchPlan(ns, ms, normalized=True, precision=precision)
        for rep in range(2):
            graphs = draw(rng, ns, ms, -1)
            say("plan normalized precision", precision, rep)
            got, _ = plan.run(graphs)
            want, _ = betweenness_batch(graphs, normalized=True, precision=precision)
            for i in range(len(ns)):
                assert_same(got[i], want[i], f"precision {precision} replay {rep} graph {i}")
        plan.close()


def test_plan_watts_strogatz_replicates():
    # Watts-Strogatz graphs of fixed sizes have fixed edge counts (n k / 2),
    # so each seed is a new replicate batch of the same shape.
    sizes = [int(x) for x in gen.ws_sizes(48, 11)]
    shape = [gen.small_world(n, 6, 0.1, 11 + i, 0) for i, n in enumerate(sizes)]
    plan = BatchPlan([e.n for e in shape], [e.m for e in shape])
    for rep in range(3):
        els = [gen.small_world(n, 6, 0.1, 1000 * rep + i, 0) for i, n in enumerate(sizes)]
        graphs = [(e.n, e.src, e.tgt) for e in els]
        say("plan ws replay", rep)
        got, _ = plan.run(graphs)
        want, _ = betweenness_batch(graphs)
        for i in range(len(els)):
            assert_same(got[i], want[i], f"ws replay {rep} graph {i}")
        e = els[rep]
        assert_same(got[rep], OracleGraph(e.n, e.src, e.tgt, None, False).betweenness(threads=0),
                    f"ws oracle {rep}")
    plan.close()


def test_plan_errors_and_recovery():
    with pytest.raises(EngineError, match=r"^graph 1: negative node or edge count$"):
        BatchPlan([3, -1], [1, 0])
    with pytest.raises(EngineError, match=r"^unknown precision code 3$"):
        BatchPlan([3], [1], precision=3)
    plan = BatchPlan([2, 3], [1, 1])
    with pytest.raises(EngineError, match=r"^graph 1: edge 0 refers to node index 5, but "
                                          r"valid indices are 0..3$"):
        plan.run([(2, [0], [1]), (3, [5], [0])])
    got, _ = plan.run([(2, [0], [1]), (3, [1], [2])])
    assert [list(x) for x in got] == [[0.0, 0.0], [0.0, 0.0, 0.0]]
    got, _ = plan.run([(2, [1], [0]), (3, [0], [2])])
    assert [list(x) for x in got] == [[0.0, 0.0], [0.0, 0.0, 0.0]]
    plan.close()
    # a path 0-1-2 through one edge set, then another
    plan = BatchPlan([3], [2])
    assert list(plan.run([(3, [0, 1], [1, 2])])[0][0]) == [0.0, 1.0, 0.0]
    assert list(plan.run([(3, [1, 0], [0, 2])])[0][0]) == [1.0, 0.0, 0.0]
    plan.close()
    # a shape with no graph for the persistent launch: every graph one at a time
    plan = BatchPlan([2600], [2599])
    got, _ = plan.run([(2600, np.arange(2599), np.arange(1, 2600))])
    assert list(got[0][:3]) == [0.0, 2598.0, 2.0 * 2597]
    plan.close()
