"""Freeman centralization (net_centralization) and centrality normalized for
two-mode data (net_twomode_centrality) on the device, against the CPU
oracle's restatement of topography.rs:376-568 and bipartite.rs:130-269 and
the committed goldens.

The bar: degree and closeness indices bit-exact against the oracle (integer
degrees, bit-exact breadth-first profiles, the same index arithmetic);
betweenness within 1e-12 relative (runs of 4,096 or more sources batch
them in a free order); the reference's tolerances against the goldens.
"""
import re
import sys

import numpy as np
import pytest

from _common import (CENTRALIZATION_VARIANTS, TWOMODE_VARIANTS, close, closeness_golden,
                     random_case)
from knappnet_amd import generate as gen
from knappnet_amd.engine import EngineError, Graph
from oracle.oracle import OracleGraph

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(a, b):
    a, b = np.atleast_1d(np.asarray(a, dtype=np.float64)), np.atleast_1d(np.asarray(b))
    return a.shape == b.shape and np.array_equal(np.isnan(a), np.isnan(b)) and \
        np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


def rel_ok(a, b, tol=1e-12):
    a, b = np.atleast_1d(np.asarray(a, dtype=np.float64)), np.atleast_1d(np.asarray(b))
    if not np.array_equal(np.isnan(a), np.isnan(b)):
        return False
    k = ~np.isnan(a)
    err = np.abs(a[k] - b[k]) / np.maximum(np.abs(b[k]), 1e-300)
    err[np.abs(a[k] - b[k]) <= 1e-14] = 0.0
    return float(err.max(initial=0.0)) <= tol


def both(n, s, t, w, directed):
    return Graph(n, s, t, w, directed), OracleGraph(n, s, t, w, directed)


def test_centralization_goldens(fixtures):
    checked = 0
    for variant, kw in CENTRALIZATION_VARIANTS.items():
        for name, fx in fixtures.items():
            if variant not in fx["golden"]:
                continue
            say("fixture", name, variant)
            g, o = both(fx["n"], fx["src"], fx["tgt"], fx["weights"], fx["directed"])
            got = g.centralization(**kw)
            want = o.centralization(**kw)
            ok = same(got, want) if kw["measure"] != 1 else rel_ok(got, want)
            assert ok, (name, variant, got, want)
            for oracle, gold in fx["golden"][variant].items():
                assert close([got], closeness_golden(gold), 1e-10, 1e-12), (name, variant, oracle)
            checked += 1
    assert checked >= 150


def test_twomode_goldens(fixtures):
    checked = 0
    for variant, kw in TWOMODE_VARIANTS.items():
        for name, fx in fixtures.items():
            if variant not in fx["golden"]:
                continue
            say("fixture", name, variant)
            g, o = both(fx["n"], fx["src"], fx["tgt"], fx["weights"], fx["directed"])
            got = g.twomode_centrality(fx["types"], **kw)
            want = o.twomode_centrality(fx["types"], **kw)
            ok = same(got, want) if kw["measure"] != 2 else rel_ok(got, want)
            assert ok, (name, variant)
            for oracle, gold in fx["golden"][variant].items():
                assert close(got, closeness_golden(gold), 1e-9, 1e-12), (name, variant, oracle)
            checked += 1
    assert checked >= 20


CENT_KW = [dict(measure=m, mode=mode, loops=lp, normalized=nz, convention=c)
           for m in (0, 1, 2) for mode in (1, 2, 3) for lp in (False, True)
           for nz in (False, True) for c in (0, 2, 3)]


@pytest.mark.parametrize("seed", range(4))
def test_centralization_random(seed):
    rng = np.random.default_rng(1300 + seed)
    for case in range(4):
        n, s, t, directed = random_case(rng, max_n=60)
        say("seed", seed, "case", case, n, len(s), directed)
        g, o = both(n, s, t, None, directed)
        for kw in CENT_KW:
            try:
                want = o.centralization(**kw)
            except Exception as e:  # the same refusal on both sides
                with pytest.raises(EngineError, match="^" + re.escape(str(e)) + "$"):
                    g.centralization(**kw)
                continue
            got = g.centralization(**kw)
            ok = same(got, want) if kw["measure"] != 1 else rel_ok(got, want)
            assert ok, (kw, got, want)


@pytest.mark.parametrize("seed", range(4))
def test_twomode_random(seed):
    rng = np.random.default_rng(1400 + seed)
    for case in range(5):
        n1, n2 = int(rng.integers(1, 30)), int(rng.integers(1, 30))
        m = int(rng.integers(0, 3 * (n1 + n2)))
        a = rng.integers(0, n1, m)
        b = rng.integers(0, n2, m) + n1
        flip = rng.random(m) < 0.5  # either direction: direction is ignored
        s, t = np.where(flip, b, a), np.where(flip, a, b)
        types = np.r_[np.zeros(n1, dtype=np.int32), np.ones(n2, dtype=np.int32)]
        perm = rng.permutation(n1 + n2)  # interleave the modes
        inv = np.argsort(perm)
        s, t, types = inv[s], inv[t], types[perm]
        directed = bool(rng.integers(0, 2))
        say("seed", seed, "case", case, n1, n2, m, directed)
        g, o = both(n1 + n2, s, t, None, directed)
        for kw in (dict(measure=0), dict(measure=0, normalized=False),
                   dict(measure=1, unreachable=1), dict(measure=1, unreachable=0),
                   dict(measure=1, unreachable=0, normalized=False), dict(measure=2),
                   dict(measure=2, normalized=False)):
            got = g.twomode_centrality(types, **kw)
            want = o.twomode_centrality(types, **kw)
            ok = same(got, want) if kw["measure"] != 2 else rel_ok(got, want)
            assert ok, (kw, got[:6], want[:6])


def test_centralization_large():
    # 4,096 nodes: betweenness and closeness in batched (free) source order
    el = gen.rmat(12, 8, directed=True, seed=9)
    g, o = both(el.n, el.src, el.tgt, None, True)
    say("rmat12", el.n, el.m)
    for kw in (dict(measure=1), dict(measure=1, convention=2), dict(measure=2, mode=1),
               dict(measure=2, mode=3, convention=2), dict(measure=2, mode=2, convention=3),
               dict(measure=0, mode=2)):
        got, want = g.centralization(**kw), o.centralization(**kw)
        ok = same(got, want) if kw["measure"] != 1 else rel_ok(got, want)
        assert ok, (kw, got, want)


def test_errors():
    g = Graph(4, [0, 1, 2], [1, 2, 3], None, True)
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
                     r'\(mode "all"\) only'),
                    ({"measure": 3}, "eigenvector centralization is not part of this library"),
                    ({"measure": 0, "convention": 3},
                     "UCINET reports degree centralization for in- and out-degree only")):
        with pytest.raises(EngineError, match="^" + msg + "$"):
            g.centralization(**kw)
    b = Graph(4, [0, 1], [2, 3], None, False)
    with pytest.raises(EngineError, match="^node types are missing$"):
        b.twomode_centrality(None)
    with pytest.raises(EngineError, match="^node 2 has type code 5; use 0 for the first mode "
                                          "and 1 for the second$"):
        b.twomode_centrality([0, 0, 5, 1])
    with pytest.raises(EngineError, match="^unknown two-mode centrality code 3$"):
        b.twomode_centrality([0, 0, 1, 1], measure=3)
    with pytest.raises(EngineError, match="^unknown code 2 for unreachable nodes$"):
        b.twomode_centrality([0, 0, 1, 1], measure=1, unreachable=2)
    with pytest.raises(EngineError, match="^edge 1 joins nodes 1 and 3, which are in the same "
                                          "mode; in a two-mode network every edge joins the two "
                                          "modes$"):
        b.twomode_centrality([0, 1, 1, 1])
    with pytest.raises(EngineError, match="^the Wasserman-Faust rule applies to normalized "
                                          "closeness only$"):
        b.twomode_centrality([0, 0, 1, 1], measure=1, unreachable=1, normalized=False)
    for n in (0, 1, 2):
        gg, oo = both(n, [], [], None, False)
        for kw in CENT_KW:
            if kw["measure"] == 0 and kw["convention"] == 3 and kw["mode"] == 3:
                continue
            assert same(gg.centralization(**kw), oo.centralization(**kw)), (n, kw)
        types = [0] * n
        for kw in (dict(measure=0), dict(measure=1), dict(measure=2)):
            assert same(gg.twomode_centrality(types, **kw), oo.twomode_centrality(types, **kw))
