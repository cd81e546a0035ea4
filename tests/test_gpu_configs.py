"""Parity at the benchmark workloads' full sizes, through the C ABI.

ER G(2000, 10000) with every node a source is covered bit for bit in
test_gpu_parity (test_er_2000_bitexact). This file runs the other four
workloads at the sizes the benchmark times them, against the CPU oracle on
samples of their sources and through properties that hold at any size:

- directed R-MAT scale 18 (n = 262,144): a strided sample of 256 sources
  against the oracle (bit-exact in exact order, within 1e-12 by default);
  levels and path counts of 8 sources bit-exact; and over every source the
  distance-sum identity sum_v b(v) = sum_s (T_s - R_s) (T_s: sum of the
  distances from s, R_s: nodes s reaches; each shortest s-t path has
  d(s, t) - 1 interior nodes), T_s and R_s taken from the device's
  closeness, a search without path counts;
- Barabasi-Albert n = 200,000 with integer lengths 1-16 (Dial): a sample
  of 128 sources against the oracle, distances and path counts of 8
  sources bit-exact, and a landmark-batched run of 4,096 sources equal
  (within 1e-12) to the sum of its two halves run in the reference's order;
- 4,096 Watts-Strogatz graphs in one netbc_betweenness_batch call: every
  64th graph bit-exact against the oracle, and 256 more through the
  distance-sum identity (from the oracle's distance profiles);
- undirected R-MAT scale 22 (n = 4,194,304) with the 8,192 sources drawn
  from seed 7: the first 64 of them against the oracle (bit-exact in exact
  order, within 1e-12 by default), levels and path counts of 4 sources
  bit-exact, the batched 8,192-source run equal to the sum of its halves,
  and the FP32-dependency form within 1e-6 of FP64 (max and mean relative
  error printed).

Tolerance rule: validation/compare.py:313-314 (|a - b| <= rel * max); the
FP64 bar is 1e-12 relative (crates/engine/src/metrics/betweenness.rs:63-115
is the arithmetic both sides follow).
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from knappnet_amd import generate as gen
from knappnet_amd.engine import Graph, betweenness_batch
from oracle.oracle import OracleGraph, available_cores

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    err = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    err[(a == 0) & (b == 0)] = 0.0
    return err


def assert_rel(a, b, what, tol=1e-12):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    err = rel_err(a, b)
    worst = float(err.max(initial=0.0))
    say(f"{what}: max rel err {worst:.3e}")
    assert worst <= tol, f"{what}: max rel err {worst:.3e} > {tol}"


def assert_same(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    if not np.array_equal(a, b):
        i = int(np.argmax(a != b))
        raise AssertionError(f"{what}: first difference at {i}: {a[i]!r} vs {b[i]!r}; "
                             f"{int(np.sum(a != b))} differ")


def check_dump(g, o, sources, weighted, what):
    """Levels (or integer distances) and path counts of each source bit for
    bit against the oracle's single-source search (betweenness.rs:161-295);
    dependencies within 1e-12. The oracle's searches run in threads (ctypes
    releases the interpreter lock)."""
    dist, sig, dl = g.state_dump(sources, weighted=weighted)
    with ThreadPoolExecutor(max_workers=min(len(sources), max(1, available_cores()))) as ex:
        res = list(ex.map(lambda s: o.single_source(int(s), weighted=weighted), sources))
    for i, s in enumerate(sources):
        _, du, _, osig, odl, _ = res[i]
        assert_same(dist[i], du, f"{what} distances from {s}")
        assert_same(sig[i], osig, f"{what} path counts from {s}")
        odl = odl.copy()
        odl[s] = 0.0  # the dump leaves the source's own dependency out
        assert_rel(dl[i], odl, f"{what} dependencies from {s}")


def distance_totals(g):
    """(T, R) per node as a source, from the device's closeness: unnormalized
    closeness is 1 / T_s and normalized closeness R_s / T_s
    (closeness.rs:131-209), NaN when s reaches nothing. T_s stays below
    2^31 here, so 1 / (1 / T_s) rounds back to T_s."""
    c = g.closeness(normalized=False)
    cn = g.closeness(normalized=True)
    reach = ~np.isnan(c)
    T = np.zeros(len(c))
    R = np.zeros(len(c))
    T[reach] = np.rint(1.0 / c[reach])
    R[reach] = np.rint(cn[reach] * T[reach])
    return T, R


# ---- directed R-MAT scale 18, every node a source ---------------------------

@pytest.fixture(scope="module")
def rmat18():
    el = gen.rmat(18, 16, directed=True, seed=1)
    say("rmat18 directed", el.n, el.m)
    assert (el.n, el.m) == (262144, 3939372)
    g = Graph(el.n, el.src, el.tgt, None, True)
    o = OracleGraph(el.n, el.src, el.tgt, None, True)
    return el, g, o


def test_rmat18_source_sample_vs_oracle(rmat18):
    el, g, o = rmat18
    sample = np.arange(0, el.n, el.n // 256)[:256]
    want = o.betweenness_sources(sample, threads=0)
    exact, _ = g.betweenness_sources(sample, exact=True)
    assert_same(exact, want, "rmat18 256 sources, exact order")
    fast, diag = g.betweenness_sources(sample)
    assert_rel(fast, want, "rmat18 256 sources, default order")
    assert diag["sigma_overflow"] == 0


def test_rmat18_levels_and_counts_bitexact(rmat18):
    el, g, o = rmat18
    sources = np.arange(0, el.n, el.n // 256)[:256][::32]
    check_dump(g, o, sources, False, "rmat18")


def test_rmat18_all_sources_distance_identity(rmat18):
    el, g, _ = rmat18
    b, diag = g.betweenness_sources()
    T, R = distance_totals(g)
    lhs = float(np.sum(b))
    rhs = float(np.sum(T - R))
    say(f"rmat18 all sources: sum b = {lhs:.17g}, sum (T - R) = {rhs:.17g}")
    assert abs(lhs - rhs) <= 1e-12 * rhs
    # every source counts itself as reached
    assert diag["reached_nodes"] == int(np.sum(R)) + el.n
    assert diag["sigma_overflow"] == 0
    # the scores are deterministic
    again, _ = g.betweenness_sources()
    assert_same(again, b, "rmat18 all sources, second run")


# ---- Barabasi-Albert n = 200,000, integer lengths 1-16 -----------------------

@pytest.fixture(scope="module")
def ba200k():
    el = gen.ba_weighted()
    say("ba200k", el.n, el.m)
    assert (el.n, el.m) == (200000, 799984)
    g = Graph(el.n, el.src, el.tgt, el.weights, False)
    o = OracleGraph(el.n, el.src, el.tgt, el.weights, False)
    assert o.path_lengths() == ("integer", 16)
    return el, g, o


def test_ba200k_source_sample_vs_oracle(ba200k):
    el, g, o = ba200k
    sample = np.arange(0, el.n, el.n // 128)[:128]
    want = o.betweenness_sources(sample, weighted=True, threads=0)
    exact, diag = g.betweenness_sources(sample, weighted=True, exact=True)
    assert diag["path_kind"] == 1
    assert_same(exact, want, "ba200k 128 sources, exact order")
    fast, _ = g.betweenness_sources(sample, weighted=True)
    assert_rel(fast, want, "ba200k 128 sources, default order")


def test_ba200k_distances_and_counts_bitexact(ba200k):
    el, g, o = ba200k
    check_dump(g, o, np.array([0, 1, 4, 999, 77777, 150000, 199998, 199999]), True, "ba200k")


def test_ba200k_batched_run_is_sum_of_halves(ba200k):
    # 4,096 sources run landmark-batched (free order); each half alone runs in
    # the reference's order. Same sums, different association.
    el, g, _ = ba200k
    src = gen.sample_sources(el.n, 4096, 5)
    whole, _ = g.betweenness_sources(src, weighted=True)
    h1, _ = g.betweenness_sources(src[:2048], weighted=True)
    h2, _ = g.betweenness_sources(src[2048:], weighted=True)
    assert_rel(whole, h1 + h2, "ba200k 4,096 sources vs halves")


# ---- 4,096 Watts-Strogatz graphs in one call ---------------------------------

def test_ws4096_one_call_vs_oracle():
    els = gen.ws_batch(4096, 11)
    assert sum(e.n for e in els) == int(np.sum(gen.ws_sizes(4096, 11)))
    parts, diag = betweenness_batch([(e.n, e.src, e.tgt) for e in els])
    say("ws4096: one call, kernel_ms", round(diag["kernel_ms"], 1))
    assert len(parts) == 4096
    for i in range(0, 4096, 64):
        e = els[i]
        want = OracleGraph(e.n, e.src, e.tgt, None, False).betweenness(threads=0)
        assert_same(parts[i], want, f"ws graph {i} (n = {e.n})")
    # 256 more graphs through the distance-sum identity: undirected, so
    # sum_v b(v) = (sum over ordered pairs of d(s, t) - 1) / 2, from the
    # oracle's distance profiles (closeness.rs:240-345)
    for i in range(5, 4096, 16):
        e = els[i]
        p = OracleGraph(e.n, e.src, e.tgt, None, False).distance_profiles(np.arange(e.n))
        want = float(np.sum(p["total"]) - np.sum(p["reached"].astype(np.float64))) / 2.0
        got = float(np.sum(parts[i]))
        assert abs(got - want) <= 1e-12 * max(want, 1.0), (i, got, want)


# ---- undirected R-MAT scale 22, 8,192 sources from seed 7 ---------------------

@pytest.fixture(scope="module")
def rmat22():
    el = gen.rmat(22, 16, directed=False, seed=1)
    src = gen.sample_sources(el.n, 8192, 7)
    say("rmat22 undirected", el.n, el.m)
    assert el.n == 1 << 22 and el.m == 64152753
    assert len(np.unique(src)) == 8192
    g = Graph(el.n, el.src, el.tgt, None, False)
    return el, src, g


@pytest.fixture(scope="module")
def rmat22_oracle(rmat22):
    el, _, _ = rmat22
    return OracleGraph(el.n, el.src, el.tgt, None, False)


def test_rmat22_first_64_sources_vs_oracle(rmat22, rmat22_oracle):
    el, src, g = rmat22
    sample = src[:64]
    want = rmat22_oracle.betweenness_sources(sample, threads=0)
    exact, _ = g.betweenness_sources(sample, exact=True)
    assert_same(exact, want, "rmat22 64 sources, exact order")
    fast, _ = g.betweenness_sources(sample)
    assert_rel(fast, want, "rmat22 64 sources, default order")


def test_rmat22_levels_and_counts_bitexact(rmat22, rmat22_oracle):
    el, src, g = rmat22
    check_dump(g, rmat22_oracle, src[[0, 1, 2, 8191]], False, "rmat22")


def test_rmat22_8192_sources_fp64_and_fp32(rmat22):
    el, src, g = rmat22
    b64, d64 = g.betweenness_sources(src)
    b32, d32 = g.betweenness_sources(src, precision=1)
    assert d64["sigma_overflow"] == 0 and d32["sigma_overflow"] == 0
    h1, _ = g.betweenness_sources(src[:4096])
    h2, _ = g.betweenness_sources(src[4096:])
    assert_rel(b64, h1 + h2, "rmat22 8,192 sources vs halves")
    nz = b64 != 0
    err = rel_err(b32[nz], b64[nz])
    say(f"rmat22 FP32 dependencies vs FP64: max rel err {err.max():.3e}, "
        f"mean {err.mean():.3e} over {int(nz.sum())} nonzero scores")
    assert np.array_equal(b32 == 0, b64 == 0)
    assert float(err.max()) <= 1e-6
