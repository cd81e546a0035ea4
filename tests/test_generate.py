"""Workload generators: Philox known answers (rng.rs:117-134), stream
behaviour, and the structural properties of each generated graph. The
reference commits no edge-list goldens, so exact edge lists are pinned by
the Philox known answers plus the fingerprints below (regression pins of
this library's own output)."""
import numpy as np

from knappnet_amd import generate as gen


def test_philox_known_answers():
    assert gen.philox([0] * 4, [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    m = 0xFFFFFFFF
    assert gen.philox([m] * 4, [m, m]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert gen.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                      [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420,
                                                    0x24126EA1]


def test_streams_are_reproducible_and_distinct():
    a = gen.stream_values(7, 3, "u64", 5)
    assert np.array_equal(a, gen.stream_values(7, 3, "u64", 5))
    assert a[0] != gen.stream_values(7, 4, "u64", 1)[0]
    assert a[0] != gen.stream_values(8, 3, "u64", 1)[0]
    # next_u64 = hi << 32 | lo of consecutive u32 words of block 0
    b0 = gen.philox([0, 0, 3, 0], [7, 0])
    assert int(a[0]) == (b0[0] << 32 | b0[1]) and int(a[1]) == (b0[2] << 32 | b0[3])
    u = gen.stream_values(7, 3, "uniform", 5)
    assert np.array_equal(u, (a >> np.uint64(11)).astype(np.float64) / 2.0**53)
    below = gen.stream_values(7, 3, "below", 1000, 10)
    assert below.min() >= 0 and below.max() <= 9


def test_permutations_are_valid():
    for n in (1, 2, 17, 1000):
        p = gen.permutation(5, 0, n)
        assert sorted(p) == list(range(n))


def test_gnm_er_2000():
    el = gen.er_2000()
    assert el.n == 2000 and el.m == 10000
    s, t = el.src.astype(np.int64), el.tgt.astype(np.int64)
    assert np.all(s < t)  # undirected pairs, smaller node first, no loops
    keys = s * 2000 + t
    assert np.all(np.diff(keys) > 0)  # sorted, distinct
    assert el.fingerprint() == ER_2000_FINGERPRINT


def test_gnm_directed_and_dense():
    el = gen.gnm(30, 700, directed=True, seed=4)
    assert el.m == 700 and np.all(el.src != el.tgt)
    keys = el.src.astype(np.int64) * 30 + el.tgt
    assert np.all(np.diff(keys) > 0)
    el = gen.gnm(10, 45, directed=False, seed=4)  # complete graph (dense branch)
    assert el.m == 45


def test_small_world():
    el = gen.small_world(200, 6, 0.1, seed=3)
    assert el.m == 600
    s, t = el.src.astype(np.int64), el.tgt.astype(np.int64)
    assert np.all(s < t)
    assert len(set(zip(s.tolist(), t.tolist()))) == 600
    el0 = gen.small_world(50, 4, 0.0, seed=3)  # no rewiring: the ring lattice
    pairs = set(zip(el0.src.tolist(), el0.tgt.tolist()))
    assert pairs == {tuple(sorted((u, (u + j) % 50))) for u in range(50) for j in (1, 2)}


def test_preferential():
    el = gen.preferential(1000, 4, seed=3)
    assert el.m == 4 * (1000 - 4)
    deg = np.bincount(np.concatenate([el.src, el.tgt]), minlength=1000)
    assert deg.min() >= 1 and deg.sum() == 2 * el.m
    assert np.all(el.src != el.tgt)


def test_ba_weighted():
    el = gen.ba_weighted(n=2000, m=4, seed=3)
    assert el.weights.min() >= 1 and el.weights.max() <= 16
    assert np.all(el.weights == np.round(el.weights))


def test_rmat_small():
    el = gen.rmat(10, 8, directed=True, seed=1, threads=3)
    s, t = el.src.astype(np.int64), el.tgt.astype(np.int64)
    assert el.n == 1024 and np.all(s != t)
    keys = s * 1024 + t
    assert np.all(np.diff(keys) > 0)
    # thread count does not change the edge list
    el2 = gen.rmat(10, 8, directed=True, seed=1, threads=1)
    assert np.array_equal(el.src, el2.src) and np.array_equal(el.tgt, el2.tgt)
    u = gen.rmat(10, 8, directed=False, seed=1)
    assert np.all(u.src < u.tgt)


def test_rmat_matches_sequential_definition():
    # Recompute a scale-6 R-MAT from Stream(seed, 0).uniform() in order.
    scale, ef, seed = 6, 4, 9
    total = (1 << scale) * ef
    u = gen.stream_values(seed, 0, "uniform", total * scale)
    a, b, c, _ = gen.RMAT_ABCD
    pairs = set()
    for e in range(total):
        r = col = 0
        for i in range(scale):
            x = u[e * scale + i]
            bit = 1 << (scale - 1 - i)
            if x < a:
                pass
            elif x < a + b:
                col |= bit
            elif x < a + b + c:
                r |= bit
            else:
                r |= bit
                col |= bit
        if r != col:
            pairs.add((r, col))
    el = gen.rmat(scale, ef, directed=True, seed=seed)
    assert list(zip(el.src.tolist(), el.tgt.tolist())) == sorted(pairs)


def test_sample_sources():
    s = gen.sample_sources(100000, 8192, 7)
    assert len(set(s.tolist())) == 8192 and s.min() >= 0 and s.max() < 100000
    assert np.array_equal(s, gen.permutation(7, 0, 100000)[:8192])


def test_ws_sizes():
    sizes = gen.ws_sizes(64, 11)
    assert sizes.min() >= 50 and sizes.max() <= 2000


def test_rmat18_and_ba_pins():
    r = gen.rmat(18, 16, directed=True, seed=1)
    assert r.m == RMAT18_EDGES and r.fingerprint() == RMAT18_FINGERPRINT
    b = gen.ba_weighted()
    assert b.m == 799_984 and b.fingerprint() == BA_FINGERPRINT


# Regression pins of this library's generated workloads (FNV-1a 64 over the
# (src, tgt) pairs as little-endian u32).
ER_2000_FINGERPRINT = 0xB09D411F0D340CB3
RMAT18_EDGES = 3_939_372
RMAT18_FINGERPRINT = 0x4DD06ED5F47D416B
BA_FINGERPRINT = 0x317F327DB8D32558
