"""Closeness and harmonic centrality on the device (net_closeness), against
the CPU oracle's restatement of closeness.rs:131-481 and the committed
topography goldens.

The bar: bit-exact against the oracle for hop counts (the per-level
bookkeeping of closeness.rs:268-293 is reproduced term for term) and for
real lengths searched by the per-thread Dijkstra; within 1e-12 relative
where the bucketed searches add each source's reciprocal (and real)
lengths in node order instead of settle order; within the reference's
tolerance (rel 1e-12, abs 1e-14, topography.yaml) of the goldens.
"""
import sys

import numpy as np
import pytest

from _common import CLOSENESS_VARIANTS, close, closeness_golden, random_case
from knappnet_amd import generate as gen
from knappnet_amd.engine import EngineError, Graph
from oracle.oracle import OracleGraph

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_same(a, b, what):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, what
    assert np.array_equal(np.isnan(a), np.isnan(b)), f"{what}: NaN pattern differs"
    keep = ~np.isnan(a)
    if not np.array_equal(a[keep], b[keep]):
        d = np.flatnonzero(keep & (a != b))
        i = int(d[0])
        raise AssertionError(f"{what}: first difference at {i}: {a[i]!r} vs {b[i]!r}; "
                             f"{len(d)} differ")


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


ALL_KW = [dict(harmonic=h, normalized=nz, unreachable=u, mode=mode)
          for mode in (1, 2, 3) for h in (False, True) for nz in (False, True)
          for u in range(4) if h or ((u != 1 or nz))]


def test_closeness_goldens(fixtures):
    checked = 0
    for variant, kw in CLOSENESS_VARIANTS.items():
        for name, fx in fixtures.items():
            if variant not in fx["golden"]:
                continue
            say("fixture", name, variant)
            g, o = both(fx["n"], fx["src"], fx["tgt"], fx["weights"], fx["directed"])
            got = g.closeness(**kw)
            assert_rel(got, o.closeness(**kw), f"{name} {variant} vs oracle", abs_=1e-14)
            for oracle, want in fx["golden"][variant].items():
                assert close(got, closeness_golden(want), 1e-12, 1e-14), (name, variant, oracle)
            checked += 1
    assert checked >= 200


@pytest.mark.parametrize("seed", range(6))
def test_closeness_random_hops_bit_exact(seed):
    rng = np.random.default_rng(900 + seed)
    for case in range(6):
        n, s, t, directed = random_case(rng, max_n=90)
        say("seed", seed, "case", case, n, len(s), directed)
        # no weights, and uniform weights (hop counts times the common length)
        for w, weighted, kind in ((None, False, "distance"),
                                  (np.full(len(s), 2.5), True, "distance"),
                                  (np.full(len(s), 0.4), True, "strength")):
            g, o = both(n, s, t, w, directed)
            for kw in ALL_KW:
                if weighted and kw["unreachable"] == 3 and not kw["harmonic"]:
                    continue
                kw = dict(kw, weighted=weighted, weights_as=kind)
                assert_same(g.closeness(**kw), o.closeness(**kw), f"{kw}")


@pytest.mark.parametrize("seed", range(6))
def test_closeness_random_weighted(seed):
    rng = np.random.default_rng(950 + seed)
    for case in range(5):
        n, s, t, directed = random_case(rng, max_n=90)
        say("seed", seed, "case", case, n, len(s), directed)
        w_int = rng.integers(1, 6, len(s)).astype(np.float64)
        w_real = rng.random(len(s)) * 3 + 0.5
        for w, kind in ((w_int, "distance"), (w_real, "distance"), (w_real, "strength")):
            g, o = both(n, s, t, w, directed)
            for kw in ALL_KW:
                if kw["unreachable"] == 3 and not kw["harmonic"]:
                    continue
                kw = dict(kw, weighted=True, weights_as=kind)
                assert_rel(g.closeness(**kw), o.closeness(**kw), f"{kw}", abs_=1e-14)


def test_closeness_very_unequal_lengths_bit_exact():
    # longest / shortest arc up to 1e5: the reference's Dijkstra per source
    rng = np.random.default_rng(12)
    for case in range(4):
        n, s, t, directed = random_case(rng, max_n=60)
        w = 10.0 ** rng.uniform(-4, 1, len(s))
        say("unequal", case, n, len(s), directed)
        g, o = both(n, s, t, w, directed)
        for mode in (1, 2, 3):
            for h in (False, True):
                kw = dict(harmonic=h, mode=mode, weighted=True, normalized=True)
                assert_same(g.closeness(**kw), o.closeness(**kw), f"{case} {kw}")


def test_closeness_deep_paths():
    # more levels than the breadth-first search keeps bitmaps for
    n = 500
    s = np.arange(n - 1)
    t = s + 1
    g, o = both(n, s, t, 1.0 + (np.arange(n - 1) % 5), True)
    for kw in (dict(mode=1), dict(mode=2, normalized=True), dict(mode=3, unreachable=3),
               dict(harmonic=True, mode=3), dict(mode=1, normalized=True, unreachable=1)):
        assert_same(g.closeness(**kw), o.closeness(**kw), f"path {kw}")
        kw = dict(kw, weighted=True)
        if kw.get("unreachable") == 3:
            continue
        assert_rel(g.closeness(**kw), o.closeness(**kw), f"path weighted {kw}")
    # a 30 x 40 grid: 68 levels from a corner
    r, c = np.divmod(np.arange(1200), 40)
    right = np.flatnonzero(c < 39)
    down = np.flatnonzero(r < 29)
    gs = np.concatenate([right, down])
    gt = np.concatenate([right + 1, down + 40])
    g, o = both(1200, gs, gt, None, False)
    for kw in (dict(normalized=True), dict(harmonic=True)):
        assert_same(g.closeness(**kw), o.closeness(**kw), f"grid {kw}")


def test_closeness_large_batched():
    # 8,192 nodes: sources batched by landmark distances, several super-batches
    el = gen.rmat(13, 8, directed=True, seed=4)
    rng = np.random.default_rng(5)
    w_int = rng.integers(1, 9, el.m).astype(np.float64)
    w_real = rng.random(el.m) * 4 + 1
    say("rmat13", el.n, el.m)
    g, o = both(el.n, el.src, el.tgt, w_int, True)
    for kw in (dict(mode=1), dict(mode=2, normalized=True, unreachable=1),
               dict(mode=3, harmonic=True), dict(mode=1, unreachable=3, normalized=True)):
        got, diag = g.closeness(options=dict(slots=8), **kw)
        assert diag["path_kind"] == 0
        assert_same(got, o.closeness(**kw), f"hops {kw}")
    for kw in (dict(mode=1, weighted=True), dict(mode=3, harmonic=True, weighted=True)):
        assert_rel(g.closeness(**kw), o.closeness(**kw), f"int {kw}")
    g, o = both(el.n, el.src, el.tgt, w_real, True)
    for kw in (dict(mode=1, weighted=True, normalized=True),
               dict(mode=2, harmonic=True, weighted=True, weights_as="strength")):
        assert_rel(g.closeness(**kw), o.closeness(**kw), f"real {kw}")


def test_closeness_errors_and_small():
    g = Graph(3, [0, 1], [1, 2], [1.0, 2.0], True)
    for kw, msg in (({"mode": 4}, "unknown degree mode code 4"),
                    ({"mode": 0}, "unknown degree mode code 0"),
                    ({"weights_as": 2}, "unknown weight kind code 2"),
                    ({"unreachable": 5}, "unknown code 5 for unreachable nodes"),
                    ({"unreachable": 1},
                     "the Wasserman-Faust closeness index is defined only in normalized form"),
                    ({"unreachable": 3, "weighted": True},
                     "closeness with unreachable lengths set to the longest length plus 1 is "
                     "defined only for unweighted networks")):
        with pytest.raises(EngineError, match="^" + msg + "$"):
            g.closeness(**kw)
    assert np.isfinite(g.closeness(harmonic=True, unreachable=1)).all()
    with pytest.raises(EngineError, match=r"^the graph has no edge weights$"):
        Graph(2, [0], [1], None, True).closeness(weighted=True)
    bad = Graph(3, [0, 1], [1, 2], [1.0, -2.0], True)
    with pytest.raises(EngineError, match=r"^edge 1 has weight -2; this algorithm treats "
                                          r"weights as distances and requires them to be "
                                          r"strictly positive$"):
        bad.closeness(weighted=True, mode=3)
    for n in (0, 1, 2):
        gg, oo = both(n, [], [], None, False)
        for kw in ALL_KW:
            assert_same(gg.closeness(**kw), oo.closeness(**kw), f"n={n} {kw}")
