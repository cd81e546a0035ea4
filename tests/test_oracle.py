"""Pins the CPU oracle (oracle/netbc_oracle.c) before it is trusted.

Checks, in order of strength:
- the reference's committed goldens for all 28 fixtures (definition oracle
  and NetworkX, with the reference's documented NetworkX differences);
- the reference's unit known answers (betweenness.rs:387-517,
  reach.rs:731-826, graph.rs:354-412, weights.rs:290-321,
  parallel.rs:644-665);
- an exact-rational evaluation of the definition (tests/definition.py) on
  random multigraphs, weighted and unweighted;
- closed forms and the distance-sum identity;
- the reference's randomised invariants (tests/properties.rs:61-250).
"""
import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import shortest_path

import definition
from _common import (CENTRALIZATION_VARIANTS, CLOSENESS_VARIANTS, EDGE_VARIANTS,
                     SUBSET_VARIANTS, TWOMODE_VARIANTS, VARIANTS, close, closeness_golden,
                     golden_values, networkx_documented, random_case, subset_call,
                     subset_networkx_documented, tolerance)
from oracle.oracle import (OracleError, OracleGraph, available_cores, ordered_sum_bits,
                           threads_for)


def og(fx):
    return OracleGraph(fx["n"], fx["src"], fx["tgt"], fx["weights"], fx["directed"])


# ---- committed goldens -------------------------------------------------------

@pytest.mark.parametrize("variant", list(VARIANTS))
def test_goldens_definition(fixtures, variant):
    rel, ab = tolerance(variant)
    checked = 0
    for name, fx in fixtures.items():
        if variant not in fx["golden"]:
            continue
        g = og(fx)
        for threads in (1, 4):
            got = g.betweenness(threads=threads, **VARIANTS[variant])
            assert close(got, fx["golden"][variant]["definition"], rel, ab), (name, variant)
        checked += 1
    assert checked >= 7


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_goldens_networkx(fixtures, variant):
    rel, ab = tolerance(variant)
    for name, fx in fixtures.items():
        if variant not in fx["golden"] or "networkx" not in fx["golden"][variant]:
            continue
        got = og(fx).betweenness(**VARIANTS[variant])
        ok = close(got, fx["golden"][variant]["networkx"], rel, ab)
        if networkx_documented(fx, variant):
            continue  # documented difference; the definition golden rules
        assert ok, (name, variant)


@pytest.mark.parametrize("variant", list(EDGE_VARIANTS))
def test_edge_goldens_definition(fixtures, variant):
    rel, ab = tolerance(variant)
    checked = 0
    for name, fx in fixtures.items():
        if variant not in fx["golden"]:
            continue
        g = og(fx)
        for threads in (1, 4):
            got = g.edge_betweenness(threads=threads, **EDGE_VARIANTS[variant])
            assert close(got, fx["golden"][variant]["definition"], rel, ab), (name, variant)
        checked += 1
    assert checked >= 7


@pytest.mark.parametrize("variant", list(EDGE_VARIANTS))
def test_edge_goldens_networkx(fixtures, variant):
    rel, ab = tolerance(variant)
    for name, fx in fixtures.items():
        if variant not in fx["golden"] or "networkx" not in fx["golden"][variant]:
            continue
        got = og(fx).edge_betweenness(**EDGE_VARIANTS[variant])
        ok = close(got, fx["golden"][variant]["networkx"], rel, ab)
        if networkx_documented(fx, variant):
            continue  # documented difference; the definition golden rules
        assert ok, (name, variant)


@pytest.mark.parametrize("variant", list(SUBSET_VARIANTS))
def test_subset_goldens(fixtures, variant):
    # centrality_extra.yaml: rel 1e-9, abs 1e-12 for every variant
    checked = 0
    for name, fx in fixtures.items():
        if variant not in fx["golden"]:
            continue
        kw, out = subset_call(fx["n"], SUBSET_VARIANTS[variant])
        nodes, edges = og(fx).betweenness_subset(**kw)
        got = nodes if out == "nodes" else edges
        for oracle, want in fx["golden"][variant].items():
            ok = close(got, golden_values(want), 1e-9, 1e-12)
            if oracle == "networkx" and subset_networkx_documented(fx, variant):
                continue
            assert ok, (name, variant, oracle)
        checked += 1
    assert checked >= 7


@pytest.mark.parametrize("variant", list(CLOSENESS_VARIANTS))
def test_closeness_goldens(fixtures, variant):
    # topography.yaml: rel 1e-12, abs 1e-14 for every variant
    checked = 0
    for name, fx in fixtures.items():
        if variant not in fx["golden"]:
            continue
        got = og(fx).closeness(**CLOSENESS_VARIANTS[variant])
        for oracle, want in fx["golden"][variant].items():
            assert close(got, closeness_golden(want), 1e-12, 1e-14), (name, variant, oracle)
        checked += 1
    assert checked >= 7


# ---- unit known answers: betweenness.rs:387-517 -------------------------------

def bc(n, s, t, w=None, directed=False, **kw):
    return OracleGraph(n, s, t, w, directed).betweenness(**kw)


def test_path_graph():
    assert list(bc(4, [0, 1, 2], [1, 2, 3])) == [0.0, 2.0, 2.0, 0.0]


def test_star_normalized_is_one():
    r = bc(5, [0, 0, 0, 0], [1, 2, 3, 4], normalized=True)
    assert abs(r[0] - 1.0) < 1e-12 and all(x == 0.0 for x in r[1:])


def test_directed_cycle():
    assert list(bc(3, [0, 1, 2], [1, 2, 0], directed=True)) == [1.0, 1.0, 1.0]


def test_split_paths_share_credit():
    assert list(bc(4, [0, 0, 1, 2], [1, 2, 3, 3])) == [0.5, 0.5, 0.5, 0.5]


def test_weights_change_paths():
    assert list(bc(3, [0, 1, 0], [1, 2, 2], [1.0, 1.0, 5.0], weighted=True)) == [0.0, 1.0, 0.0]


def test_tied_float_paths_are_detected():
    r = bc(4, [0, 1, 0, 2], [1, 3, 2, 3], [0.1, 0.2, 0.15, 0.15], directed=True, weighted=True)
    assert abs(r[1] - 0.5) < 1e-12 and abs(r[2] - 0.5) < 1e-12


def test_rejects_non_positive_weights():
    g = OracleGraph(2, [0], [1], [0.0], True)
    with pytest.raises(OracleError, match=r"^edge 0 has weight 0; this algorithm treats "
                                          r"weights as distances and requires them to be "
                                          r"strictly positive$"):
        g.betweenness(weighted=True)
    g = OracleGraph(3, [0, 1], [1, 2], [2.5, -1e-7], True)
    with pytest.raises(OracleError, match=r"^edge 1 has weight -0.0000001;"):
        g.betweenness(weighted=True)


def test_every_algorithm_agrees_on_integer_weights():
    s, t = [0, 0, 1, 2, 1, 3, 4], [1, 2, 3, 3, 4, 5, 5]
    w = [2.0, 1.0, 1.0, 2.0, 3.0, 2.0, 1.0]
    g = OracleGraph(6, s, t, w, False)
    dial = g.betweenness(weighted=True, algorithm="dial")
    assert np.array_equal(dial, g.betweenness(weighted=True, algorithm="dijkstra"))
    assert np.array_equal(dial, g.betweenness(weighted=True, algorithm="auto"))


def test_strengths_are_inverted():
    g = OracleGraph(3, [0, 1, 0], [1, 2, 2], [1.0, 1.0, 4.0], False)
    assert list(g.betweenness(weighted=True, weights_as="strength")) == [0.0, 0.0, 0.0]
    assert list(g.betweenness(weighted=True)) == [0.0, 1.0, 0.0]


# ---- edge betweenness known answers: reach.rs:731-826 ---------------------------

def ebc(n, s, t, w=None, directed=False, **kw):
    return OracleGraph(n, s, t, w, directed).edge_betweenness(**kw)


def test_edge_betweenness_by_hand():
    assert list(ebc(3, [0, 1], [1, 2])) == [2.0, 2.0]
    assert list(ebc(3, [0, 1], [1, 2], directed=True)) == [2.0, 2.0]
    # parallel edges split the paths; a loop carries nothing
    assert list(ebc(3, [0, 0, 1, 2], [1, 1, 2, 2])) == [1.0, 1.0, 2.0, 0.0]
    sq = ([0, 1, 2, 3], [1, 2, 3, 0])
    assert list(ebc(4, *sq)) == [2.0] * 4
    assert close(ebc(4, *sq, normalized=True), [2.0 / 6.0] * 4, 1e-15, 1e-15)
    # normalisation is undefined below two nodes
    assert np.isnan(ebc(1, [0], [0], normalized=True)).all()
    assert ebc(0, [], []).size == 0


def test_edge_betweenness_sums_to_total_path_length():
    e = [(0, 1), (0, 2), (1, 2), (2, 3), (3, 4), (4, 5), (3, 5), (1, 5)]
    s, t = zip(*e)
    eb = ebc(6, s, t)
    D = shortest_path(csr_matrix((np.ones(len(e)), (s, t)), shape=(6, 6)), directed=False,
                      unweighted=True)
    assert abs(eb.sum() - D[np.triu_indices(6, 1)].sum()) < 1e-12


def test_weighted_edge_flow_conservation():
    # Into every node v flows b(v) + (nodes reaching v): each path through
    # v enters it once, and each path ending at v enters it once.
    e = [(0, 1), (0, 2), (1, 3), (2, 3), (1, 4), (3, 5), (4, 5)]
    w = [2.0, 1.0, 1.0, 2.0, 3.0, 2.0, 1.0]
    s, t = zip(*e)
    g = OracleGraph(6, s, t, w, True)
    eb = g.edge_betweenness(weighted=True)
    nb = g.betweenness(weighted=True)
    D = shortest_path(csr_matrix((w, (s, t)), shape=(6, 6)), directed=True)
    reached = (np.isfinite(D) & (D > 0)).sum(axis=0)
    inflow = np.zeros(6)
    np.add.at(inflow, np.asarray(t), eb)
    assert close(inflow, nb + reached, 1e-12, 1e-12)


# ---- graph build: graph.rs:354-412 -------------------------------------------

def rows(off, tgt, v):
    return list(tgt[off[v]:off[v + 1]])


def test_directed_rows_are_sorted_and_complete():
    g = OracleGraph(4, [2, 0, 0, 1, 0], [1, 3, 1, 2, 1], None, True)
    off, tgt, _ = g.csr()
    assert rows(off, tgt, 0) == [1, 1, 3]
    assert rows(off, tgt, 1) == [2]
    assert rows(off, tgt, 2) == [1]
    off, tgt, _ = g.csr(incoming=True)
    assert rows(off, tgt, 1) == [0, 0, 2]
    assert rows(off, tgt, 3) == [0]


def test_undirected_stores_both_directions_and_loops_once():
    g = OracleGraph(3, [0, 1, 2], [1, 2, 2], [1.0, 2.0, 3.0], False)
    off, tgt, w = g.csr()
    assert rows(off, tgt, 1) == [0, 2] and list(w[off[1]:off[2]]) == [1.0, 2.0]
    assert rows(off, tgt, 2) == [1, 2] and list(w[off[2]:off[3]]) == [2.0, 3.0]
    assert g.arc_count() == 5


def test_parallel_arcs_keep_input_order():
    g = OracleGraph(3, [0, 0, 0, 0], [2, 1, 2, 1], [5.0, 6.0, 7.0, 8.0], True)
    off, tgt, w = g.csr()
    assert rows(off, tgt, 0) == [1, 1, 2, 2]
    assert list(w[:4]) == [6.0, 8.0, 5.0, 7.0]


def test_rejects_bad_input():
    with pytest.raises(OracleError, match=r"^edge 0 refers to node index 2, but valid "
                                          r"indices are 0..2$"):
        OracleGraph(2, [0], [2], None, True)
    with pytest.raises(OracleError, match=r"^edge 0 has a missing, NaN, or infinite weight$"):
        OracleGraph(2, [0], [1], [float("nan")], True)
    with pytest.raises(OracleError, match=r"^negative node or edge count$"):
        OracleGraph(-1, [], [], None, True)


# ---- path lengths: weights.rs:290-321 -----------------------------------------

def chain(w):
    n = len(w) + 1
    return OracleGraph(n, list(range(len(w))), list(range(1, n)), w, True)


def test_chooses_the_cheapest_exact_algorithm():
    assert chain([2.5, 2.5]).path_lengths() == ("uniform", 0)
    assert chain([1.0, 3.0]).path_lengths() == ("integer", 3)
    assert chain([1.0, 5000.0]).path_lengths()[0] == "real"
    assert chain([0.5, 2.0]).path_lengths()[0] == "real"
    assert chain([0.5, 2.0]).path_lengths(weighted=False) == ("uniform", 0)
    assert chain([1.0, 0.5]).path_lengths(weights_as="strength") == ("integer", 2)
    with pytest.raises(OracleError):
        chain([0.5]).path_lengths(algorithm="dial")
    with pytest.raises(OracleError):
        chain([0.0]).path_lengths(weights_as="strength")
    with pytest.raises(OracleError, match=r"^the graph has no edge weights$"):
        OracleGraph(2, [0], [1], None, True).path_lengths()


# ---- parallel.rs:644-665 -------------------------------------------------------

def test_threads_follow_the_work():
    G = 1 << 17
    assert threads_for(8, 100, 10) == 1
    assert threads_for(8, 100, 3 * G) == 3
    assert threads_for(8, 2, 100 * G) == 2
    assert threads_for(8, 100, 2**64 - 1) == min(8, 100)
    assert threads_for(1, 100, 2**64 - 1) == 1
    assert threads_for(0, 100, 2**64 - 1) == min(available_cores(), 100)


def test_sums_in_order_for_every_thread_count():
    for items in (0, 1, 5, 63, 64, 65, 1000, 4097):
        one = ordered_sum_bits(items, 1)
        for t in (2, 3, 8, 100):
            assert ordered_sum_bits(items, t) == one, (items, t)


# ---- exact definition on random graphs ----------------------------------------

@pytest.mark.parametrize("seed", range(40))
def test_definition_random(seed):
    rng = np.random.default_rng(1000 + seed)
    n, s, t, directed = random_case(rng, max_n=14)
    w = rng.integers(1, 5, size=len(s)).astype(np.float64)
    g = OracleGraph(n, s, t, w, directed)
    for normalized in (False, True):
        want = definition.betweenness(n, s, t, w, directed, normalized=normalized)
        assert close(g.betweenness(normalized=normalized), want, 1e-12, 1e-12)
    want = definition.betweenness(n, s, t, w, directed, weighted=True)
    assert close(g.betweenness(weighted=True), want, 1e-12, 1e-12)
    want = definition.betweenness(n, s, t, w, directed, weighted=True, strength=True)
    assert close(g.betweenness(weighted=True, weights_as="strength"), want, 1e-9, 1e-9)


@pytest.mark.parametrize("seed", range(30))
def test_edge_definition_random(seed):
    rng = np.random.default_rng(2000 + seed)
    n, s, t, directed = random_case(rng, max_n=12)
    w = rng.integers(1, 5, size=len(s)).astype(np.float64)
    g = OracleGraph(n, s, t, w, directed)
    for normalized in (False, True):
        want = definition.edge_betweenness(n, s, t, w, directed, normalized=normalized)
        assert close(g.edge_betweenness(normalized=normalized), want, 1e-12, 1e-12)
    want = definition.edge_betweenness(n, s, t, w, directed, weighted=True)
    assert close(g.edge_betweenness(weighted=True), want, 1e-12, 1e-12)
    want = definition.edge_betweenness(n, s, t, w, directed, weighted=True, strength=True)
    assert close(g.edge_betweenness(weighted=True, weights_as="strength"), want, 1e-9, 1e-9)
    # results never depend on the thread count
    assert np.array_equal(g.edge_betweenness(threads=1), g.edge_betweenness(threads=3))


@pytest.mark.parametrize("seed", range(24))
def test_subset_definition_random(seed):
    # sources with repeats, a target set, endpoints, parallel edges distinct
    # or collapsed, unit and integer lengths, against exact rationals
    rng = np.random.default_rng(3000 + seed)
    n, s, t, directed = random_case(rng, max_n=11)
    w = rng.integers(1, 4, size=len(s)).astype(np.float64)
    srcs = list(rng.integers(0, n, size=int(rng.integers(1, n + 2))))
    tg = list(rng.choice(n, size=int(rng.integers(1, n + 1)), replace=False))
    half = 1.0 if directed else 0.5
    g = OracleGraph(n, s, t, w, directed)
    for collapse in (0, 1):
        for weighted in (False, True):
            for endpoints in (False, True):
                wn, we = definition.subset_raw(n, s, t, w, directed, weighted=weighted,
                                               sources=srcs, targets=tg, endpoints=endpoints,
                                               collapse=bool(collapse))
                gn, ge = g.betweenness_subset(sources=srcs, targets=tg, weighted=weighted,
                                              endpoints=endpoints, parallel=collapse)
                assert close(gn, np.array(wn) * half, 1e-12, 1e-12)
                assert close(ge, np.array(we) * half, 1e-12, 1e-12)


def test_subset_with_every_pair_is_betweenness():
    # path_centrality.rs:32-35: S = T = every node, unordered pairs, distinct
    # parallel edges give betweenness and edge betweenness
    rng = np.random.default_rng(12)
    for _ in range(20):
        n, s, t, directed = random_case(rng, max_n=30)
        w = rng.random(len(s)) + 0.5
        g = OracleGraph(n, s, t, w, directed)
        for kw in ({}, {"weighted": True}):
            nodes, edges = g.betweenness_subset(**kw)
            assert close(nodes, g.betweenness(**kw), 1e-12, 1e-12)
            assert close(edges, g.edge_betweenness(**kw), 1e-12, 1e-12)


def test_subset_errors():
    g = OracleGraph(3, [0, 1], [1, 2], [1.0, -2.0], True)
    with pytest.raises(OracleError, match=r"^node index 5 is out of range for a network "
                                          r"with 3 nodes$"):
        g.betweenness_subset(targets=[0, 5])
    for kw, msg in (({"rescale": 3}, "unknown rescaling code 3"),
                    ({"parallel": 2}, "unknown code 2 for parallel edges"),
                    ({"edge_split": -1}, "unknown edge split code -1"),
                    ({"tie": 2}, "unknown tie code 2"),
                    ({"weights_as": 4}, "unknown weight kind code 4")):
        with pytest.raises(OracleError, match="^" + msg + "$"):
            g.betweenness_subset(**kw)
    with pytest.raises(OracleError, match=r"^edge 1 has length -2; shortest-path betweenness "
                                          r"needs strictly positive lengths$"):
        g.betweenness_subset(weighted=True)
    with pytest.raises(OracleError, match=r"^edge 1 has weight -2; weights read as strengths "
                                          r"become lengths 1/w and must be strictly positive$"):
        g.betweenness_subset(weighted=True, weights_as="strength")
    with pytest.raises(OracleError, match=r"^the graph has no edge weights$"):
        OracleGraph(2, [0], [1], None, True).betweenness_subset(weighted=True)


@pytest.mark.parametrize("seed", range(12))
def test_closeness_definition_random(seed):
    # closeness and harmonic centrality from scipy's all-pairs distances
    rng = np.random.default_rng(4000 + seed)
    n, s, t, directed = random_case(rng, max_n=40)
    w = rng.random(len(s)) * 3 + 0.5
    for weighted in (False, True):
        lens = w if weighted else np.ones(len(s))
        A = csr_matrix((np.full(len(s), np.inf), (s, t)), shape=(n, n)).tolil()
        for e in range(len(s)):  # the lightest of parallel edges, loops ignored
            if s[e] != t[e] and lens[e] < A[s[e], t[e]]:
                A[s[e], t[e]] = lens[e]
        A = A.tocsr()
        A.data[np.isinf(A.data)] = 0
        A.eliminate_zeros()
        for mode in (1, 2, 3):
            D = shortest_path(A, directed=directed and mode != 3)
            if directed and mode == 2:
                D = D.T
            off = ~np.eye(n, dtype=bool)
            fin = np.isfinite(D) & off
            r = fin.sum(axis=1)
            tot = np.where(fin, D, 0).sum(axis=1)
            harm = np.where(fin, 1.0 / np.where(fin, D, 1.0), 0).sum(axis=1)
            g = OracleGraph(n, s, t, w, directed)
            kw = dict(mode=mode, weighted=weighted)
            with np.errstate(divide="ignore", invalid="ignore"):
                want = np.where(r > 0, r / tot, np.nan)
            assert close(g.closeness(normalized=True, **kw), want, 1e-12, 1e-14)
            assert close(g.closeness(harmonic=True, **kw), harm, 1e-12, 1e-14)


def test_closeness_errors():
    g = OracleGraph(3, [0, 1], [1, 2], [1.0, 2.0], True)
    for kw, msg in (({"mode": 4}, "unknown degree mode code 4"),
                    ({"weights_as": 2}, "unknown weight kind code 2"),
                    ({"unreachable": 5}, "unknown code 5 for unreachable nodes"),
                    ({"unreachable": 1},
                     "the Wasserman-Faust closeness index is defined only in normalized form"),
                    ({"unreachable": 3, "weighted": True},
                     "closeness with unreachable lengths set to the longest length plus 1 is "
                     "defined only for unweighted networks")):
        with pytest.raises(OracleError, match="^" + msg + "$"):
            g.closeness(**kw)
    assert np.isfinite(g.closeness(harmonic=True, unreachable=1)).all()


# ---- closed forms and the distance-sum identity -------------------------------

def test_closed_forms():
    for n in (2, 3, 7, 50):
        r = bc(n, list(range(n - 1)), list(range(1, n)))
        assert list(r) == [float(i * (n - 1 - i)) for i in range(n)]
    n = 9
    r = bc(n, [0] * (n - 1), list(range(1, n)))
    assert r[0] == (n - 1) * (n - 2) / 2 and all(x == 0 for x in r[1:])
    s, t = zip(*[(i, j) for i in range(6) for j in range(i + 1, 6)])
    assert all(x == 0 for x in bc(6, s, t))
    n = 11
    r = bc(n, list(range(n)), [(i + 1) % n for i in range(n)], directed=True)
    assert all(x == (n - 1) * (n - 2) / 2 for x in r)


@pytest.mark.parametrize("seed", range(10))
def test_distance_sum_identity(seed):
    # sum_v b(v) = sum over ordered reachable pairs s != t of (d(s,t) - 1),
    # halved for undirected graphs (each interior node of each shortest
    # path gets the path's share; shares of one pair sum to d - 1).
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 200))
    m = int(rng.integers(n, 4 * n))
    s = rng.integers(0, n, m)
    t = rng.integers(0, n, m)
    directed = bool(seed % 2)
    g = OracleGraph(n, s, t, None, directed)
    A = csr_matrix((np.ones(m), (s, t)), shape=(n, n))
    D = shortest_path(A, directed=directed, unweighted=True)
    fin = np.isfinite(D) & (D > 0)
    total = float(np.sum(D[fin] - 1))
    if not directed:
        total /= 2
    assert close([g.betweenness().sum()], [total], 1e-12, 1e-9)


# ---- randomised invariants: crates/engine/tests/properties.rs ----------------

CASES = 60


def test_relabeling_nodes_permutes_results():
    rng = np.random.default_rng(1)
    for _ in range(CASES):
        n, s, t, directed = random_case(rng)
        perm = rng.permutation(n)
        a = bc(n, s, t, directed=directed)
        b = bc(n, perm[s], perm[t], directed=directed)
        assert close(a, b[perm])


def test_unit_weights_match_unweighted():
    rng = np.random.default_rng(2)
    for _ in range(CASES):
        n, s, t, directed = random_case(rng)
        g = OracleGraph(n, s, t, np.ones(len(s)), directed)
        assert close(g.betweenness(), g.betweenness(weighted=True))


def test_results_do_not_depend_on_thread_count():
    rng = np.random.default_rng(3)
    for _ in range(CASES):
        n, s, t, directed = random_case(rng)
        g = OracleGraph(n, s, t, None, directed)
        assert np.array_equal(g.betweenness(threads=1), g.betweenness(threads=4))


def test_undirected_betweenness_matches_symmetric_directed():
    rng = np.random.default_rng(5)
    for _ in range(CASES):
        n, s, t, _ = random_case(rng)
        keep = s != t
        s2 = np.concatenate([s, t[keep]])
        t2 = np.concatenate([t, s[keep]])
        bu = bc(n, s, t)
        bd = bc(n, s2, t2, directed=True) / 2
        assert close(bu, bd)


def test_every_path_algorithm_agrees():
    rng = np.random.default_rng(6)
    for _ in range(CASES):
        n, s, t, directed = random_case(rng)
        w = (1 + rng.integers(0, 6, len(s))).astype(np.float64)
        g = OracleGraph(n, s, t, w, directed)
        dial = g.betweenness(weighted=True, algorithm="dial")
        heap = g.betweenness(weighted=True, algorithm="dijkstra")
        assert close(dial, heap)
        assert close(dial, g.betweenness(weighted=True))
        h = OracleGraph(n, s, t, 1.0 / w, directed)
        assert close(g.betweenness(weighted=True, weights_as="strength"),
                     h.betweenness(weighted=True))


def test_uniform_weights_match_unweighted_under_every_algorithm():
    rng = np.random.default_rng(7)
    for _ in range(CASES):
        n, s, t, directed = random_case(rng)
        k = float(1 + rng.integers(0, 9))
        g = OracleGraph(n, s, t, np.full(len(s), k), directed)
        plain = g.betweenness()
        for alg in ("auto", "dial", "dijkstra"):
            assert close(plain, g.betweenness(weighted=True, algorithm=alg))


def test_full_source_list_is_ordinary_betweenness():
    # path_centrality.rs:955-997: the full source list gives betweenness().
    rng = np.random.default_rng(8)
    for _ in range(20):
        n, s, t, directed = random_case(rng)
        g = OracleGraph(n, s, t, None, directed)
        assert np.array_equal(g.betweenness_sources(np.arange(n)), g.betweenness())


# ---- Freeman centralization and two-mode centrality ---------------------------

@pytest.mark.parametrize("variant", list(CENTRALIZATION_VARIANTS))
def test_centralization_goldens(fixtures, variant):
    # topography.yaml: rel 1e-12 / 1e-10, abs 1e-14 / 1e-12; NaN written as -1
    checked = 0
    for name, fx in fixtures.items():
        if variant not in fx["golden"]:
            continue
        got = og(fx).centralization(**CENTRALIZATION_VARIANTS[variant])
        for oracle, want in fx["golden"][variant].items():
            assert close([got], closeness_golden(want), 1e-10, 1e-12), (name, variant, oracle)
        checked += 1
    assert checked >= 8


@pytest.mark.parametrize("variant", list(TWOMODE_VARIANTS))
def test_twomode_goldens(fixtures, variant):
    # twomode.yaml: rel 1e-14 (degree), 1e-12 (closeness), 1e-9 (betweenness)
    checked = 0
    for name, fx in fixtures.items():
        if variant not in fx["golden"]:
            continue
        got = og(fx).twomode_centrality(fx["types"], **TWOMODE_VARIANTS[variant])
        for oracle, want in fx["golden"][variant].items():
            assert close(got, closeness_golden(want), 1e-9, 1e-12), (name, variant, oracle)
        checked += 1
    assert checked >= 3


def test_centralization_stars():
    # a star reaches the theoretical maximum of every index: 1
    for n in (3, 5, 9):
        s, t = np.zeros(n - 1, dtype=np.int64), np.arange(1, n)
        g = OracleGraph(n, s, t, None, False)
        for kw in (dict(measure=0, loops=False), dict(measure=0, convention=2),
                   dict(measure=1), dict(measure=1, convention=3), dict(measure=2),
                   dict(measure=2, convention=2), dict(measure=2, convention=3)):
            assert abs(g.centralization(**kw) - 1.0) < 1e-14, (n, kw)
        # igraph with loops counted: the maximum is (n-1) n, the star (n-1)(n-2)
        assert abs(g.centralization(measure=0) - (n - 2) / n) < 1e-14
        # raw index: n * max - sum
        assert g.centralization(measure=0, normalized=False, loops=False) == \
            n * (n - 1) - 2 * (n - 1)
        d = OracleGraph(n, s, t, None, True)  # out-star
        assert abs(d.centralization(measure=0, mode=1, loops=False) - 1.0) < 1e-14
        assert abs(d.centralization(measure=0, mode=1, convention=3) - 1.0) < 1e-14


def test_twomode_betweenness_max_is_attained():
    # Exhaustive search over small bipartite graphs: the largest normalized
    # betweenness of a first-mode node is exactly 1, so the maximum of
    # bipartite.rs:228-241 is attained and never exceeded.
    import itertools
    for own, other in ((2, 2), (3, 2), (2, 3), (4, 2), (3, 3), (1, 3), (4, 1)):
        n = own + other
        types = [0] * own + [1] * other
        pairs = [(a, own + b) for a in range(own) for b in range(other)]
        best = 0.0
        for mask in range(1, 1 << len(pairs)):
            es = [pairs[i] for i in range(len(pairs)) if mask >> i & 1]
            g = OracleGraph(n, [e[0] for e in es], [e[1] for e in es], None, False)
            b = g.twomode_centrality(types, measure=2)
            if not np.isnan(b[0]):
                best = max(best, float(np.nanmax(b[:own])))
        want = 1.0 if other > 1 or own == 1 else 0.0
        assert abs(best - want) < 1e-12, (own, other, best)


def test_centralization_and_twomode_errors():
    g = OracleGraph(4, [0, 1, 2], [1, 2, 3], None, True)
    for kw, msg in (({"measure": 4}, "unknown centralization measure code 4"),
                    ({"mode": 0}, "unknown degree mode code 0"),
                    ({"convention": 7}, "unknown convention code 7"),
                    ({"convention": 1}, "NetworkX has no centralization index; use the igraph, "
                                        "sna, or UCINET convention"),
                    ({"measure": 3, "convention": 2},
                     "eigenvector centralization is provided for the igraph convention only, "
                     "not sna"),
                    ({"measure": 3, "mode": 1},
                     'eigenvector centralization is provided with direction ignored '
                     '\\(mode "all"\\) only'),
                    ({"measure": 3}, "eigenvector centralization is not part of this library"),
                    ({"measure": 0, "convention": 3},
                     "UCINET reports degree centralization for in- and out-degree only")):
        with pytest.raises(OracleError, match="^" + msg + "$"):
            g.centralization(**kw)
    b = OracleGraph(4, [0, 1], [2, 3], None, False)
    with pytest.raises(OracleError, match="^node types are missing$"):
        b.twomode_centrality(None)
    with pytest.raises(OracleError, match="^node 2 has type code 5; use 0 for the first mode "
                                          "and 1 for the second$"):
        b.twomode_centrality([0, 0, 5, 1])
    with pytest.raises(OracleError, match="^unknown two-mode centrality code 3$"):
        b.twomode_centrality([0, 0, 1, 1], measure=3)
    with pytest.raises(OracleError, match="^unknown code 2 for unreachable nodes$"):
        b.twomode_centrality([0, 0, 1, 1], measure=1, unreachable=2)
    with pytest.raises(OracleError, match="^edge 1 joins nodes 1 and 3, which are in the same "
                                          "mode; in a two-mode network every edge joins the two "
                                          "modes$"):
        b.twomode_centrality([0, 1, 1, 1])
    with pytest.raises(OracleError, match="^the Wasserman-Faust rule applies to normalized "
                                          "closeness only$"):
        b.twomode_centrality([0, 0, 1, 1], measure=1, unreachable=1, normalized=False)
