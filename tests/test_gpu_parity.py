"""Parity of the HIP path with the CPU oracle, through the C ABI.

The bar: device CSR, levels/distances and sigma bit-exact; FP64 scores
bit-exact against the oracle's restatement of the reference (same operation
order; the reference's own tolerance for the committed goldens is rel 1e-9,
abs 1e-9 / 1e-12 normalised). FP32-dependency runs: max relative error
<= 1e-6 against FP64, stated in the test.
"""
import sys

import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import breadth_first_order, connected_components

from _common import VARIANTS, close, random_case, tolerance
from knappnet_amd import generate as gen
from knappnet_amd.engine import EngineError, Graph
from oracle.oracle import OracleGraph

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(n, s, t, w=None, directed=False):
    return Graph(n, s, t, w, directed), OracleGraph(n, s, t, w, directed)


def assert_rel(a, b, what, tol=1e-12):
    """FP64 scores with heavy rows summed in chunks: within 1e-12 relative."""
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    err = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    err[(a == 0) & (b == 0)] = 0.0
    assert float(err.max(initial=0.0)) <= tol, f"{what}: max rel err {err.max():.3e}"


def assert_same(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    if not np.array_equal(a, b):
        i = int(np.argmax(a != b))
        raise AssertionError(f"{what}: first difference at {i}: {a[i]!r} vs {b[i]!r}; "
                             f"{int(np.sum(a != b))} differ")


def test_device_info():
    from knappnet_amd import device_info
    info = device_info()
    say("device:", info)
    assert "gfx" in info


@pytest.mark.parametrize("directed", [False, True])
def test_device_csr_matches_host_build(directed):
    rng = np.random.default_rng(11)
    for case in range(10):
        n = int(rng.integers(1, 3000))
        m = int(rng.integers(0, 5 * n))
        s = rng.integers(0, n, m)
        t = rng.integers(0, n, m)
        w = rng.random(m) + 0.5
        say("csr case", case, n, m, directed)
        g, o = pair(n, s, t, w, directed)
        for incoming in (False, True):
            off, tgt, ww = g.csr(incoming)
            ooff, otgt, ow = o.csr(incoming)
            assert_same(off.astype(np.uint64), ooff, "offsets")
            assert_same(tgt, otgt, "targets")
            assert_same(ww, ow, "weights")


def test_fixtures_bitexact_and_goldens(fixtures):
    for name, fx in fixtures.items():
        g, o = pair(fx["n"], fx["src"], fx["tgt"], fx["weights"], fx["directed"])
        for variant, spec in VARIANTS.items():
            if variant not in fx["golden"]:
                continue
            say("fixture", name, variant)
            got = g.betweenness(**spec)
            want = o.betweenness(**spec)
            assert_same(got, want, f"{name} {variant}")
            rel, ab = tolerance(variant)
            assert close(got, fx["golden"][variant]["definition"], rel, ab), (name, variant)


def test_reference_known_answers():
    cases = [
        (dict(n=4, s=[0, 1, 2], t=[1, 2, 3]), {}, [0.0, 2.0, 2.0, 0.0]),
        (dict(n=3, s=[0, 1, 2], t=[1, 2, 0], directed=True), {}, [1.0, 1.0, 1.0]),
        (dict(n=4, s=[0, 0, 1, 2], t=[1, 2, 3, 3]), {}, [0.5] * 4),
        (dict(n=3, s=[0, 1, 0], t=[1, 2, 2], w=[1.0, 1.0, 5.0]), {"weighted": True},
         [0.0, 1.0, 0.0]),
        (dict(n=3, s=[0, 1, 0], t=[1, 2, 2], w=[1.0, 1.0, 4.0]),
         {"weighted": True, "weights_as": "strength"}, [0.0, 0.0, 0.0]),
    ]
    for graph, kw, want in cases:
        g = Graph(graph["n"], graph["s"], graph["t"], graph.get("w"),
                  graph.get("directed", False))
        assert list(g.betweenness(**kw)) == want
    g = Graph(5, [0, 0, 0, 0], [1, 2, 3, 4])
    r = g.betweenness(normalized=True)
    assert abs(r[0] - 1.0) < 1e-12 and all(x == 0.0 for x in r[1:])
    g = Graph(4, [0, 1, 0, 2], [1, 3, 2, 3], [0.1, 0.2, 0.15, 0.15], directed=True)
    r = g.betweenness(weighted=True)
    assert abs(r[1] - 0.5) < 1e-12 and abs(r[2] - 0.5) < 1e-12


def test_device_error_paths():
    g = Graph(2, [0], [1], [0.0], True)
    with pytest.raises(EngineError, match=r"^edge 0 has weight 0; this algorithm treats "
                                          r"weights as distances and requires them to be "
                                          r"strictly positive$"):
        g.betweenness(weighted=True)
    g = Graph(3, [0, 1], [1, 2], [2.5, -1e-7], True)
    with pytest.raises(EngineError, match=r"^edge 1 has weight -0.0000001;"):
        g.betweenness(weighted=True, weights_as="strength")
    g = Graph(0, [], [], None, False)
    assert len(g.betweenness()) == 0
    g = Graph(3, [], [], None, True)
    assert list(g.betweenness(normalized=True)) == [0.0, 0.0, 0.0]


@pytest.mark.parametrize("seed", range(6))
def test_random_multigraphs_bitexact(seed):
    rng = np.random.default_rng(500 + seed)
    for case in range(12):
        n, s, t, directed = random_case(rng, max_n=60)
        w = (1 + rng.integers(0, 9, len(s))).astype(np.float64)
        say("random", seed, case, n, len(s), directed)
        g, o = pair(n, s, t, w, directed)
        for kw in ({}, {"normalized": True}, {"weighted": True},
                   {"weighted": True, "weights_as": "strength"}):
            assert_same(g.betweenness(**kw), o.betweenness(**kw), f"{seed}/{case} {kw}")


def test_real_lengths_bitexact():
    rng = np.random.default_rng(3)
    for case in range(8):
        n, s, t, directed = random_case(rng, max_n=80)
        w = rng.random(len(s)) * 3.5 + 0.5
        say("real", case, n, len(s), directed)
        g, o = pair(n, s, t, w, directed)
        assert_same(g.betweenness(weighted=True), o.betweenness(weighted=True), "real")


def test_er_2000_bitexact():
    el = gen.er_2000()
    say("er_2000")
    g, o = pair(el.n, el.src, el.tgt, None, False)
    got, diag = g.betweenness_sources()
    want = o.betweenness(threads=0)
    assert_same(got, want, "er_2000")
    assert_same(g.betweenness(normalized=True), o.betweenness(normalized=True, threads=0),
                "er_2000 normalized")
    assert diag["sigma_overflow"] == 0
    # reached nodes / arcs: every node's component, summed over sources
    A = csr_matrix((np.ones(el.m), (el.src, el.tgt)), shape=(el.n, el.n))
    A = A + A.T
    deg = np.diff(A.indptr)
    ncomp, lab = connected_components(A, directed=False)
    size = np.bincount(lab)
    arcs = np.bincount(lab, weights=deg)
    assert diag["reached_nodes"] == int(np.sum(size[lab]))
    assert diag["reached_arcs"] == int(np.sum(arcs[lab]))


@pytest.mark.parametrize("directed", [True, False])
def test_rmat_bitexact(directed):
    # R-MAT scale 12 has rows longer than 1024 arcs: the exact mode sums them
    # in CSR order (bit-exact); the default splits them into chunks
    # (deterministic, within 1e-12).
    el = gen.rmat(12, 32, directed=directed, seed=5)
    deg = np.bincount(el.src, minlength=el.n) + (0 if directed else
                                                 np.bincount(el.tgt, minlength=el.n))
    say("rmat12", directed, el.m, "max degree", int(deg.max()))
    assert deg.max() > 1024
    g, o = pair(el.n, el.src, el.tgt, None, directed)
    want = o.betweenness(threads=0)
    exact, _ = g.betweenness_sources(exact=True)
    assert_same(exact, want, "rmat12 exact")
    fast = g.betweenness()
    assert_rel(fast, want, "rmat12 split")
    again, _ = g.betweenness_sources()
    assert_same(again, fast, "split mode is deterministic")
    # levels and path counts stay bit-exact in either mode
    srcs = np.array([0, 1, 2, 17, 4095])
    dist, sig, _ = g.state_dump(srcs)
    for i, src in enumerate(srcs):
        _, du, _, osig, _, _ = o.single_source(src)
        assert_same(dist[i], du, f"dist {src}")
        assert_same(sig[i], osig, f"sigma {src}")


@pytest.mark.parametrize("directed", [True, False])
def test_search_directions_bitexact(directed):
    # Top-down only (1), bottom-up at every level (2) and the automatic
    # choice (0) give the oracle's scores bit for bit. R-MAT scale 12 has
    # hubs (in-degree > 512) and heavy rows (out-degree > 1024).
    el = gen.rmat(12, 32, directed=directed, seed=9)
    g, o = pair(el.n, el.src, el.tgt, None, directed)
    want = o.betweenness(threads=0)
    for direction in (1, 2, 0):
        say("directions", directed, direction)
        got, diag = g.betweenness_sources(exact=True, direction=direction)
        assert_same(got, want, f"direction {direction}")
        if direction == 1:
            assert diag["bottom_up_levels"] == 0
        else:
            assert diag["bottom_up_levels"] > 0
    # partial last batch and repeated sources, bottom-up throughout
    srcs = np.concatenate([gen.sample_sources(el.n, 150, 3), [7, 7, 7, 4095]])
    want = o.betweenness_sources(srcs, threads=0)
    for direction in (2, 0):
        got, _ = g.betweenness_sources(srcs, exact=True, direction=direction)
        assert_same(got, want, f"source list, direction {direction}")


@pytest.mark.parametrize("directed", [True, False])
def test_deep_graphs_bitexact(directed):
    # A path of 400 nodes with short random chords (and 20 random edges):
    # depth 37 undirected, 104 directed, beyond the 14 level bitmaps kept
    # from the forward search, so deeper ones alternate and are rebuilt
    # during the sweep; many tied shortest paths.
    rng = np.random.default_rng(17 + directed)
    n = 400
    a = rng.integers(0, n - 4, 300)
    src = np.concatenate([np.arange(n - 1), a, rng.integers(0, n, 20)])
    tgt = np.concatenate([np.arange(1, n), a + rng.integers(2, 4, 300), rng.integers(0, n, 20)])
    g, o = pair(n, src, tgt, None, directed)
    want = o.betweenness(threads=0)
    for direction in (1, 2, 0):
        say("deep", directed, direction)
        got, diag = g.betweenness_sources(exact=True, direction=direction)
        assert diag["levels"] > 32
        assert_same(got, want, f"deep direction {direction}")
    got, _ = g.betweenness_sources(np.arange(n).repeat(12)[: 4096 + 100])
    assert_rel(got, o.betweenness_sources(np.arange(n).repeat(12)[: 4096 + 100], threads=0),
               "deep, neighbourhood batches")


def test_search_directions_random_multigraphs():
    rng = np.random.default_rng(4242)
    for case in range(12):
        n, s, t, directed = random_case(rng, max_n=300)
        g, o = pair(n, s, t, None, directed)
        want = o.betweenness(threads=0)
        for direction in (1, 2):
            got, _ = g.betweenness_sources(exact=True, direction=direction)
            assert_same(got, want, f"case {case} direction {direction}")


def test_dial_ba_bitexact():
    el = gen.ba_weighted(n=3000, m=4, seed=5)
    say("ba3000", el.m)
    g, o = pair(el.n, el.src, el.tgt, el.weights, False)
    got, diag = g.betweenness_sources(weighted=True)
    assert diag["path_kind"] == 1
    assert_same(got, o.betweenness(weighted=True, threads=0), "ba3000")


def test_dial_long_rows_first():
    # nodes with more than 512 arcs are listed at the end of each bucket's
    # list and stepped first; two hubs of 1,400 and 900 arcs on a BA graph,
    # undirected and directed, bit-exact in the reference's order
    el = gen.ba_weighted(n=3000, m=4, seed=8)
    rng = np.random.default_rng(9)
    hub_a = rng.choice(np.arange(2, el.n), 1400, replace=False)
    hub_b = rng.choice(np.arange(2, el.n), 900, replace=False)
    s = np.concatenate([el.src, np.zeros(1400, np.int64), np.ones(900, np.int64)])
    t = np.concatenate([el.tgt, hub_a, hub_b])
    w = np.concatenate([el.weights, rng.integers(1, 17, 2300).astype(np.float64)])
    for directed in (False, True):
        say("ba3000 + hubs, directed", directed)
        g, o = pair(el.n, s, t, w, directed)
        got, diag = g.betweenness_sources(weighted=True, exact=True)
        assert diag["path_kind"] == 1
        assert_same(got, o.betweenness(weighted=True, threads=0), f"hubs directed={directed}")


def test_dial_free_order_weighted_landmarks():
    # 5,000 sources: batched by weighted distances to the landmarks (integer
    # lengths, undirected); only the association of the FP64 sums changes
    el = gen.ba_weighted(n=5000, m=4, seed=11)
    say("ba5000", el.m)
    g, o = pair(el.n, el.src, el.tgt, el.weights, False)
    got, diag = g.betweenness_sources(weighted=True)
    assert diag["path_kind"] == 1
    assert_rel(got, o.betweenness(weighted=True, threads=0), "ba5000 free order")
    exact, _ = g.betweenness_sources(weighted=True, exact=True)
    assert_same(exact, o.betweenness(weighted=True, threads=0), "ba5000 exact order")
    src = np.concatenate([np.arange(el.n), np.arange(0, el.n, 3)])
    got, _ = g.betweenness_sources(src, weighted=True)
    assert_rel(got, o.betweenness_sources(src, weighted=True, threads=0), "ba5000 repeats")
    # real lengths (bucketed): keys in units of the shortest arc
    w = np.random.default_rng(6).random(el.m) * 7 + 1
    g, o = pair(el.n, el.src, el.tgt, w, False)
    got, diag = g.betweenness_sources(weighted=True)
    assert diag["path_kind"] == 2
    assert_rel(got, o.betweenness(weighted=True, threads=0), "ba5000 real free order")
    # directed: distances to the landmarks along out-arcs
    el = gen.rmat(12, 8, directed=True, seed=21)
    w = np.random.default_rng(4).integers(1, 12, el.m).astype(np.float64)
    g, o = pair(el.n, el.src, el.tgt, w, True)
    say("rmat12 directed integer", el.m)
    assert_rel(g.betweenness(weighted=True), o.betweenness(weighted=True, threads=0),
               "rmat12 directed free order")


@pytest.mark.parametrize("weighted", [False, True])
def test_state_dump_bitexact(weighted):
    rng = np.random.default_rng(77)
    for case in range(6):
        n, s, t, directed = random_case(rng, max_n=120)
        w = (1 + rng.integers(0, 7, len(s))).astype(np.float64)
        g, o = pair(n, s, t, w, directed)
        srcs = np.arange(n)
        say("dump", weighted, case, n, len(s), directed)
        dist, sig, dl = g.state_dump(srcs, weighted=weighted)
        for src in srcs:
            kind, du, _, osig, odl, _ = o.single_source(src, weighted=weighted)
            assert_same(dist[src], du, f"dist {case}/{src}")
            assert_same(sig[src], osig, f"sigma {case}/{src}")
            odl = odl.copy()
            odl[src] = 0.0
            assert_same(dl[src], odl, f"delta {case}/{src}")


def test_slots_do_not_change_bits():
    el = gen.rmat(11, 8, directed=False, seed=2)
    g = Graph(el.n, el.src, el.tgt, None, False)
    base, _ = g.betweenness_sources(slots=1)
    for slots in (2, 3, 7, 32):
        say("slots", slots)
        got, d = g.betweenness_sources(slots=slots)
        assert_same(got, base, f"slots={slots}")


def test_source_list_matches_oracle():
    el = gen.rmat(11, 8, directed=True, seed=3)
    srcs = gen.sample_sources(el.n, 300, 7)
    g, o = pair(el.n, el.src, el.tgt, None, True)
    say("source list")
    want = o.betweenness_sources(srcs, threads=0)
    got, _ = g.betweenness_sources(srcs, exact=True)
    assert_same(got, want, "sources")
    got, _ = g.betweenness_sources(srcs)
    assert_rel(got, want, "sources split")
    full, _ = g.betweenness_sources(np.arange(el.n))
    assert_same(full, g.betweenness(), "full list")


def test_default_mode_large_source_sets():
    # 4096 or more sources: batches by neighbourhood (and chunked heavy rows)
    # in the default mode, reference order in the exact mode.
    for directed in (True, False):
        el = gen.rmat(13, 8, directed=directed, seed=21)
        g, o = pair(el.n, el.src, el.tgt, None, directed)
        say("default mode", directed, el.n, el.m)
        want = o.betweenness(threads=0)
        exact, _ = g.betweenness_sources(exact=True)
        assert_same(exact, want, "exact")
        fast, _ = g.betweenness_sources()
        assert_rel(fast, want, "default")
        assert_same(g.betweenness_sources()[0], fast, "deterministic")
        assert_rel(g.betweenness(normalized=True), o.betweenness(normalized=True, threads=0),
                   "normalized")


def test_every_node_a_hub_free_order():
    # 4,096 nodes, each with 70 in-arcs (every node is a hub): all counts
    # come from the hub tasks, whose rows must be cleared between
    # super-batches (8 slots: 8 super-batches); scores are the sweep's
    # per-(slot, node) sums folded in slot order.
    rng = np.random.default_rng(77)
    n, k = 4096, 70
    t = np.repeat(np.arange(n), k)
    s = rng.integers(0, n - 1, n * k)
    s = np.where(s >= t, s + 1, s)  # no self-loops
    g, o = pair(n, s, t, None, True)
    say("all hubs", n, len(s))
    want = o.betweenness(threads=0)
    got, diag = g.betweenness_sources(slots=8)
    assert_rel(got, want, "slots=8")
    # the fold also counts reached pairs: every source reaches every node
    adj = csr_matrix((np.ones(len(s)), (s, t)), shape=(n, n))
    assert connected_components(adj, directed=True, connection="strong")[0] == 1
    assert diag["reached_nodes"] == n * n and diag["reached_arcs"] == n * len(s)
    assert_same(g.betweenness_sources(slots=8)[0], got, "deterministic")
    assert_rel(g.betweenness_sources()[0], want, "64 slots")
    assert_rel(g.edge_betweenness_sources(slots=8)[0], o.edge_betweenness(threads=0), "edges")


def test_fp32_dependencies_within_1e6():
    # FP32 coefficients with Kahan-compensated successor sums; sigma and the
    # cross-source sum in FP64. Tolerance: max relative error <= 1e-6
    # against the FP64 result on nodes with non-zero score.
    for directed, scale in ((False, 12), (True, 12)):
        el = gen.rmat(scale, 16, directed=directed, seed=9)
        g = Graph(el.n, el.src, el.tgt, None, directed)
        say("fp32", directed)
        f64, _ = g.betweenness_sources(precision=0)
        f32, _ = g.betweenness_sources(precision=1)
        nz = f64 > 0
        rel = np.abs(f32[nz] - f64[nz]) / f64[nz]
        say("fp32 max rel err", float(rel.max()), "mean", float(rel.mean()))
        assert rel.max() <= 1e-6
        assert np.all(f32[~nz] == 0)


def test_sigma_overflow_is_flagged():
    # 60 diamonds in series: 2^60 shortest paths end to end.
    k = 60
    s, t = [], []
    for i in range(k):
        a, b, c, d = 3 * i, 3 * i + 1, 3 * i + 2, 3 * i + 3
        s += [a, a, b, c]
        t += [b, c, d, d]
    n = 3 * k + 1
    g = Graph(n, s, t, None, True)
    _, diag = g.betweenness_sources()
    assert diag["sigma_overflow"] == 1 and diag["max_sigma"] == 2.0**60
    g = Graph(n, s[: 4 * 20], t[: 4 * 20], None, True)
    _, diag = g.betweenness_sources()
    assert diag["sigma_overflow"] == 0 and diag["max_sigma"] == 2.0**20


@pytest.mark.parametrize("directed", [False, True])
def test_batch_of_graphs_bitexact(directed):
    from knappnet_amd.engine import betweenness_batch
    rng = np.random.default_rng(31)
    graphs = []
    for i in range(40):
        n = int(rng.integers(1, 300)) if i % 7 else int(rng.integers(1500, 2100))
        m = int(rng.integers(0, 4 * n))
        graphs.append((n, rng.integers(0, n, m), rng.integers(0, n, m)))
    graphs.append((0, [], []))
    graphs.append((2500, rng.integers(0, 2500, 5000), rng.integers(0, 2500, 5000)))
    say("batch", directed, len(graphs))
    for normalized in (False, True):
        parts, diag = betweenness_batch(graphs, directed=directed, normalized=normalized)
        for i, (n, s, t) in enumerate(graphs):
            want = OracleGraph(n, s, t, None, directed).betweenness(normalized=normalized,
                                                                    threads=0)
            assert_same(parts[i], want, f"graph {i} (n={n})")


def test_batch_watts_strogatz_and_weighted():
    from knappnet_amd.engine import betweenness_batch
    sizes = gen.ws_sizes(24, 11)
    els = [gen.small_world(int(sizes[i]), 6, 0.1, 11 + i, 0) for i in range(24)]
    graphs = [(e.n, e.src, e.tgt) for e in els]
    say("ws batch")
    parts, diag = betweenness_batch(graphs)
    for i, e in enumerate(els):
        assert_same(parts[i], OracleGraph(e.n, e.src, e.tgt, None, False).betweenness(threads=0),
                    f"ws {i}")
    # a deep graph (path of 600 nodes: 599 levels) goes to the level-synchronous engine
    deep = (600, np.arange(599), np.arange(1, 600))
    parts, _ = betweenness_batch([deep, graphs[0]])
    assert list(parts[0]) == [float(i * (599 - i)) for i in range(600)]
    rng = np.random.default_rng(4)
    wg = [(e.n, e.src, e.tgt, 1.0 + rng.integers(0, 16, e.m)) for e in els[:4]]
    parts, _ = betweenness_batch(wg, weighted=True)
    for i, (n, s, t, w) in enumerate(wg):
        assert_same(parts[i], OracleGraph(n, s, t, w, False).betweenness(weighted=True,
                                                                          threads=0), f"w {i}")
    with pytest.raises(EngineError, match=r"^graph 1: edge 0 refers to node index 5, but "
                                          r"valid indices are 0..3$"):
        betweenness_batch([(2, [0], [1]), (3, [5], [0])])


def test_runs_in_sequence_on_one_handle():
    # Count rows are left in place between super-batches and runs (only hub
    # rows are cleared): every kind of run on one handle, in turn, must
    # still match the oracle.
    el = gen.rmat(13, 8, directed=True, seed=31)
    w = np.random.default_rng(8).integers(1, 9, el.m).astype(np.float64)
    g, o = pair(el.n, el.src, el.tgt, w, True)
    want = o.betweenness(threads=0)
    say("rmat13 sequence", el.n, el.m)
    for rnd in range(2):
        assert_rel(g.betweenness(), want, f"free order {rnd}")
        exact, _ = g.betweenness_sources(exact=True)
        assert_same(exact, want, f"exact order {rnd}")
        assert_rel(g.edge_betweenness(), o.edge_betweenness(threads=0), f"edge {rnd}")
        nodes, _ = g.betweenness_subset(targets=np.arange(0, el.n, 3))
        assert_rel(nodes, o.betweenness_subset(targets=np.arange(0, el.n, 3))[0], f"subset {rnd}")
        c = g.closeness(normalized=True)
        assert np.array_equal(c, o.closeness(normalized=True), equal_nan=True), f"closeness {rnd}"
        assert_rel(g.betweenness(weighted=True), o.betweenness(weighted=True, threads=0),
                   f"weighted {rnd}")
        f32, _ = g.betweenness_sources(precision=1)
        assert_rel(f32, want, f"fp32 {rnd}", tol=1e-6)
