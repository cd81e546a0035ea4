"""Seeded synthetic workloads for the betweenness path.

Edge lists come from ``libnetgen.so`` (Philox4x32-10 streams named by
(seed, stream index), crates/engine/src/rng.rs:22-110). The five named
workloads:

- ``er_2000``: G(n=2,000, m=10,000), undirected, no loops, seed 0
  (generate.rs:255-290, replicate 0 = Stream(0, 0)).
- ``rmat(18, directed=True)``: R-MAT scale 18, edge factor 16,
  a,b,c,d = 0.57,0.19,0.19,0.05, seed 1, loops dropped, duplicates removed,
  isolates kept (see ``gen_rmat`` in csrc/generate.c for the definition).
- ``ba_weighted``: Barabasi-Albert n=200,000, 4 edges per new node
  (generate.rs:602-631) from Stream(seed, 0), with integer distance
  weights 1 + Stream(seed, 1).below(16) per edge in edge order, as FP64.
- ``ws_batch``: 4,096 Watts-Strogatz graphs, graph i with
  n_i = 50 + Stream(base, 2^32 + i).below(1951), k = 6, p = 0.1, drawn from
  Stream(base + i, 0) (generate.rs:533-581).
- ``rmat(22, directed=False)`` with 8,192 sources: the first 8,192 entries
  of Stream(7, 0).permutation(n) (rng.rs:101-109).
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_int64, c_uint32, c_uint64
from dataclasses import dataclass

import numpy as np

from . import _native

RMAT_ABCD = (0.57, 0.19, 0.19, 0.05)


@dataclass
class EdgeList:
    n: int
    src: np.ndarray  # uint32
    tgt: np.ndarray  # uint32
    directed: bool
    weights: np.ndarray | None = None

    @property
    def m(self) -> int:
        return len(self.src)

    def fingerprint(self) -> int:
        return fingerprint(self.src, self.tgt)


def _u32p(a):
    return a.ctypes.data_as(POINTER(c_uint32))


def fingerprint(src: np.ndarray, tgt: np.ndarray) -> int:
    lib = _native.generators()
    s = np.ascontiguousarray(src, dtype=np.uint32)
    t = np.ascontiguousarray(tgt, dtype=np.uint32)
    return int(lib.gen_fingerprint(_u32p(s), _u32p(t), len(s)))


def philox(counter, key):
    lib = _native.generators()
    c = np.asarray(counter, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib.gen_philox4x32_10(_u32p(c), _u32p(k), _u32p(o))
    return [int(x) for x in o]


def stream_values(seed: int, index: int, kind: str, count: int, n: int = 0) -> np.ndarray:
    lib = _native.generators()
    out = np.zeros(count, dtype=np.uint64)
    code = {"u64": 0, "below": 1, "uniform": 2}[kind]
    lib.gen_stream_values(seed, index, code, n, count, out.ctypes.data_as(POINTER(c_uint64)))
    return out.view(np.float64) if kind == "uniform" else out


def permutation(seed: int, index: int, n: int) -> np.ndarray:
    lib = _native.generators()
    out = np.zeros(n, dtype=np.int64)
    lib.gen_permutation(seed, index, n, out.ctypes.data_as(POINTER(c_int64)))
    return out


def gnm(n: int, m: int, directed: bool = False, loops: bool = False, seed: int = 0,
        rep: int = 0) -> EdgeList:
    lib = _native.generators()
    s = np.zeros(max(m, 1), dtype=np.uint32)
    t = np.zeros(max(m, 1), dtype=np.uint32)
    got = lib.gen_gnm(n, m, int(directed), int(loops), seed, rep, _u32p(s), _u32p(t))
    if got < 0:
        raise ValueError("too many edges for G(n, m)")
    return EdgeList(n, s[:got], t[:got], directed)


def small_world(n: int, k: int, p: float, seed: int, index: int = 0) -> EdgeList:
    lib = _native.generators()
    m = n * k // 2
    s = np.zeros(max(m, 1), dtype=np.uint32)
    t = np.zeros(max(m, 1), dtype=np.uint32)
    got = lib.gen_small_world(n, k, p, seed, index, _u32p(s), _u32p(t))
    if got < 0:
        raise ValueError("invalid Watts-Strogatz parameters")
    return EdgeList(n, s[:got], t[:got], False)


def preferential(n: int, m: int, seed: int, index: int = 0) -> EdgeList:
    lib = _native.generators()
    cnt = m * (n - m)
    s = np.zeros(max(cnt, 1), dtype=np.uint32)
    t = np.zeros(max(cnt, 1), dtype=np.uint32)
    got = lib.gen_preferential(n, m, seed, index, _u32p(s), _u32p(t))
    if got < 0:
        raise ValueError("invalid preferential-attachment parameters")
    return EdgeList(n, s[:got], t[:got], False)


def rmat(scale: int, edge_factor: int = 16, directed: bool = True, seed: int = 1,
         abcd=RMAT_ABCD, threads: int = 16) -> EdgeList:
    lib = _native.generators()
    fp = POINTER(c_uint32)()
    tp = POINTER(c_uint32)()
    a, b, c, _ = abcd
    m = lib.gen_rmat(scale, edge_factor, a, b, c, int(directed), seed, threads,
                     ctypes.byref(fp), ctypes.byref(tp))
    if m < 0:
        raise ValueError("invalid R-MAT parameters")
    s = np.ctypeslib.as_array(fp, shape=(max(m, 1),))[:m].copy()
    t = np.ctypeslib.as_array(tp, shape=(max(m, 1),))[:m].copy()
    lib.gen_free(ctypes.cast(fp, ctypes.c_void_p))
    lib.gen_free(ctypes.cast(tp, ctypes.c_void_p))
    return EdgeList(1 << scale, s, t, directed)


def er_2000() -> EdgeList:
    return gnm(2000, 10000, directed=False, loops=False, seed=0, rep=0)


def ba_weighted(n: int = 200_000, m: int = 4, seed: int = 3, wmax: int = 16) -> EdgeList:
    el = preferential(n, m, seed, 0)
    w = stream_values(seed, 1, "below", el.m, wmax).astype(np.float64) + 1.0
    el.weights = w
    return el


def ba_real(n: int = 200_000, m: int = 4, seed: int = 3, lo: float = 1.0,
            hi: float = 16.0) -> EdgeList:
    """Barabasi-Albert graph with real weights uniform in [lo, hi) (stream 2)."""
    el = preferential(n, m, seed, 0)
    el.weights = lo + (hi - lo) * stream_values(seed, 2, "uniform", el.m)
    return el


def ws_sizes(count: int = 4096, base: int = 11) -> np.ndarray:
    lib = _native.generators()
    out = np.zeros(count, dtype=np.int64)
    one = np.zeros(1, dtype=np.uint64)
    for i in range(count):
        lib.gen_stream_values(base, (1 << 32) + i, 1, 1951, 1,
                              one.ctypes.data_as(POINTER(c_uint64)))
        out[i] = 50 + int(one[0])
    return out


def ws_batch(count: int = 4096, base: int = 11, k: int = 6, p: float = 0.1) -> list[EdgeList]:
    sizes = ws_sizes(count, base)
    return [small_world(int(sizes[i]), k, p, base + i, 0) for i in range(count)]


def sample_sources(n: int, count: int = 8192, seed: int = 7) -> np.ndarray:
    return permutation(seed, 0, n)[:count].copy()
