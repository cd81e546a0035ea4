"""ctypes wrapper of the CPU oracle (oracle/netbc_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the benchmark's
CPU-baseline leg, as the checker. The product library never imports it.
Build with ``make -C oracle`` (``__graft_entry__.build()`` does it).
"""
from __future__ import annotations

import ctypes
from ctypes import (POINTER, c_char_p, c_double, c_int32, c_int64, c_size_t,
                    c_uint32, c_uint64, c_void_p)
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libnetbc_oracle.so"
ERR_LEN = 1024
_lib = None


class OracleError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"oracle not built ({LIB}); run `make -C oracle`")
        L = ctypes.CDLL(str(LIB))
        L.orc_graph_new.argtypes = [c_int64, POINTER(c_int64), POINTER(c_int64),
                                    POINTER(c_double), c_int64, c_int32, POINTER(c_void_p),
                                    c_char_p, c_size_t]
        L.orc_graph_new.restype = c_int32
        L.orc_graph_free.argtypes = [c_void_p]
        L.orc_graph_free.restype = None
        L.orc_graph_arc_count.argtypes = [c_void_p, c_int32]
        L.orc_graph_arc_count.restype = c_int64
        L.orc_graph_csr.argtypes = [c_void_p, c_int32, POINTER(c_uint64), POINTER(c_uint32),
                                    POINTER(c_double)]
        L.orc_graph_csr.restype = None
        L.orc_path_lengths.argtypes = [c_void_p, c_int32, c_int32, c_int32, POINTER(c_int32),
                                       POINTER(c_uint32), c_char_p, c_size_t]
        L.orc_path_lengths.restype = c_int32
        L.orc_betweenness_alg.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_int32, c_int32,
                                          POINTER(c_double), c_char_p, c_size_t]
        L.orc_betweenness_alg.restype = c_int32
        L.orc_edge_betweenness.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_int32,
                                           POINTER(c_double), ctypes.c_char_p, ctypes.c_size_t]
        L.orc_edge_betweenness.restype = c_int32
        L.orc_edge_betweenness_sources.argtypes = [c_void_p, POINTER(c_int64), c_int64, c_int32,
                                                   c_int32, c_int32, c_int32, POINTER(c_double),
                                                   c_char_p, c_size_t]
        L.orc_edge_betweenness_sources.restype = c_int32
        L.orc_betweenness_sources.argtypes = [c_void_p, POINTER(c_int64), c_int64, c_int32,
                                              c_int32, c_int32, c_int32, POINTER(c_double),
                                              c_char_p, c_size_t]
        L.orc_betweenness_sources.restype = c_int32
        L.orc_betweenness_subset.argtypes = [c_void_p, POINTER(c_int64), c_int64,
                                             POINTER(c_int64), c_int64, c_int32, c_int32,
                                             c_int32, c_int32, c_int32, c_int32, c_int32,
                                             c_int32, POINTER(c_double), POINTER(c_double),
                                             c_char_p, c_size_t]
        L.orc_betweenness_subset.restype = c_int32
        L.orc_closeness.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_int32, c_int32,
                                    c_int32, POINTER(c_double), c_char_p, c_size_t]
        L.orc_closeness.restype = c_int32
        L.orc_single_source.argtypes = [c_void_p, c_int32, c_int32, c_int64, POINTER(c_int64),
                                        POINTER(c_int32), POINTER(c_uint64), POINTER(c_double),
                                        POINTER(c_double), POINTER(c_double), c_char_p,
                                        c_size_t]
        L.orc_single_source.restype = c_int32
        L.orc_ordered_sum_test.argtypes = [c_uint64, c_uint64]
        L.orc_ordered_sum_test.restype = c_double
        L.orc_threads_for.argtypes = [c_int64, c_int64, c_uint64]
        L.orc_threads_for.restype = c_int64
        L.orc_distance_profiles.argtypes = [
            c_void_p, c_int32, c_int32, POINTER(c_int64), c_int64, c_int32, POINTER(c_uint64),
            POINTER(c_double), POINTER(c_double), POINTER(c_double), c_char_p, c_size_t]
        L.orc_distance_profiles.restype = c_int32
        L.orc_centralization.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_int32, c_int32,
                                         POINTER(c_double), c_char_p, c_size_t]
        L.orc_centralization.restype = c_int32
        L.orc_twomode_centrality.argtypes = [c_void_p, POINTER(c_int32), c_int32, c_int32,
                                             c_int32, POINTER(c_double), c_char_p, c_size_t]
        L.orc_twomode_centrality.restype = c_int32
        L.orc_available_cores.argtypes = []
        L.orc_available_cores.restype = c_uint64
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(POINTER(t))


ALG = {"auto": 0, "bfs": 1, "dial": 2, "dijkstra": 3}
KIND = {"distance": 0, "strength": 1}


class OracleGraph:
    def __init__(self, n, src, tgt, weights=None, directed=False):
        self.L = lib()
        self.n = int(n)
        s = np.ascontiguousarray(np.asarray(src, dtype=np.int64))
        t = np.ascontiguousarray(np.asarray(tgt, dtype=np.int64))
        w = None if weights is None else np.ascontiguousarray(np.asarray(weights, np.float64))
        self.h = c_void_p()
        err = ctypes.create_string_buffer(ERR_LEN)
        if self.L.orc_graph_new(self.n, _p(s, c_int64), _p(t, c_int64),
                                None if w is None else _p(w, c_double), len(s),
                                int(bool(directed)), ctypes.byref(self.h), err, ERR_LEN):
            self.h = c_void_p()
            raise OracleError(err.value.decode())
        self.directed = bool(directed)
        self.weighted = w is not None
        self.m = len(s)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_graph_free(self.h)
            self.h = c_void_p()

    def arc_count(self, incoming=False) -> int:
        return int(self.L.orc_graph_arc_count(self.h, int(incoming)))

    def csr(self, incoming=False):
        A = self.arc_count(incoming)
        off = np.zeros(self.n + 1, dtype=np.uint64)
        tgt = np.zeros(max(A, 1), dtype=np.uint32)
        w = np.zeros(max(A, 1), dtype=np.float64) if self.weighted else None
        self.L.orc_graph_csr(self.h, int(incoming), _p(off, c_uint64), _p(tgt, c_uint32),
                             None if w is None else _p(w, c_double))
        return off, tgt[:A], None if w is None else w[:A]

    def path_lengths(self, weighted=True, weights_as="distance", algorithm="auto"):
        kind = c_int32()
        mx = c_uint32()
        err = ctypes.create_string_buffer(ERR_LEN)
        if self.L.orc_path_lengths(self.h, int(weighted), KIND[weights_as], ALG[algorithm],
                                   ctypes.byref(kind), ctypes.byref(mx), err, ERR_LEN):
            raise OracleError(err.value.decode())
        return ["uniform", "integer", "real"][kind.value], mx.value

    def betweenness(self, weighted=False, normalized=False, threads=1, weights_as="distance",
                    algorithm="auto") -> np.ndarray:
        out = np.zeros(max(self.n, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        kind = KIND[weights_as] if isinstance(weights_as, str) else int(weights_as)
        if self.L.orc_betweenness_alg(self.h, int(weighted), kind, int(normalized),
                                      int(threads), ALG[algorithm], _p(out, c_double), err,
                                      ERR_LEN):
            raise OracleError(err.value.decode())
        return out[: self.n]

    def edge_betweenness(self, weighted=False, normalized=False, threads=1,
                         weights_as="distance") -> np.ndarray:
        """reach.rs:174-232: one value per input edge."""
        out = np.zeros(max(self.m, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        if self.L.orc_edge_betweenness(self.h, int(weighted), KIND[weights_as], int(normalized),
                                       int(threads), _p(out, c_double), err, ERR_LEN):
            raise OracleError(err.value.decode())
        return out[: self.m]

    def edge_betweenness_sources(self, sources, weighted=False, normalized=False, threads=1,
                                 weights_as="distance") -> np.ndarray:
        s = np.ascontiguousarray(np.asarray(sources, dtype=np.int64))
        out = np.zeros(max(self.m, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        if self.L.orc_edge_betweenness_sources(self.h, _p(s, c_int64), len(s), int(weighted),
                                               KIND[weights_as], int(normalized), int(threads),
                                               _p(out, c_double), err, ERR_LEN):
            raise OracleError(err.value.decode())
        return out[: self.m]

    def betweenness_sources(self, sources, weighted=False, normalized=False, threads=1,
                            weights_as="distance") -> np.ndarray:
        s = np.ascontiguousarray(np.asarray(sources, dtype=np.int64))
        out = np.zeros(max(self.n, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        if self.L.orc_betweenness_sources(self.h, _p(s, c_int64), len(s), int(weighted),
                                          KIND[weights_as], int(normalized), int(threads),
                                          _p(out, c_double), err, ERR_LEN):
            raise OracleError(err.value.decode())
        return out[: self.n]

    def betweenness_subset(self, sources=None, targets=None, weighted=False,
                           weights_as="distance", endpoints=False, normalized=False, rescale=0,
                           parallel=0, edge_split=0, tie=0):
        """path_centrality.rs:421-593: (node values, edge values). sources /
        targets None: every node."""
        def lst(x):
            if x is None:
                return None, -1
            a = np.ascontiguousarray(np.asarray(x, dtype=np.int64))
            return a, len(a)
        s, ns = lst(sources)
        t, nt = lst(targets)
        nodes = np.zeros(max(self.n, 1), dtype=np.float64)
        edges = np.zeros(max(self.m, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        kind = KIND[weights_as] if isinstance(weights_as, str) else int(weights_as)
        if self.L.orc_betweenness_subset(
                self.h, None if s is None else _p(s, c_int64), ns,
                None if t is None else _p(t, c_int64), nt, int(weighted), kind, int(endpoints),
                int(normalized), int(rescale), int(parallel), int(edge_split), int(tie),
                _p(nodes, c_double), _p(edges, c_double), err, ERR_LEN):
            raise OracleError(err.value.decode())
        return nodes[: self.n], edges[: self.m]

    def closeness(self, harmonic=False, mode=1, weighted=False, weights_as="distance",
                  normalized=False, unreachable=0):
        """closeness.rs:131-209 (mode 1 out, 2 in, 3 all; unreachable 0 ignore,
        1 Wasserman-Faust, 2 zero, 3 max plus one)."""
        out = np.zeros(max(self.n, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        kind = KIND[weights_as] if isinstance(weights_as, str) else int(weights_as)
        if self.L.orc_closeness(self.h, int(harmonic), int(mode), int(weighted), kind,
                                int(normalized), int(unreachable), _p(out, c_double), err,
                                ERR_LEN):
            raise OracleError(err.value.decode())
        return out[: self.n]

    def centralization(self, measure=0, mode=3, loops=True, normalized=True, convention=0):
        """topography.rs:376-568: measure 0 degree, 1 betweenness, 2 closeness;
        convention 0 igraph, 2 sna, 3 UCINET."""
        out = np.zeros(1, dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        if self.L.orc_centralization(self.h, int(measure), int(mode), int(bool(loops)),
                                     int(bool(normalized)), int(convention), _p(out, c_double),
                                     err, ERR_LEN):
            raise OracleError(err.value.decode())
        return float(out[0])

    def twomode_centrality(self, types, measure=0, normalized=True, unreachable=0):
        """bipartite.rs:130-269: measure 0 degree, 1 closeness, 2 betweenness."""
        out = np.zeros(max(self.n, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        tp = None
        if types is not None:
            ta = np.ascontiguousarray(types, dtype=np.int32)
            tp = _p(ta, c_int32)
        if self.L.orc_twomode_centrality(self.h, tp, int(measure), int(bool(normalized)),
                                         int(unreachable), _p(out, c_double), err, ERR_LEN):
            raise OracleError(err.value.decode())
        return out[: self.n]

    def distance_profiles(self, sources, weighted=False, weights_as="distance", threads=0):
        """closeness.rs:240-345 for the listed sources (out-paths): a dict of
        reached, total, harmonic and longest arrays."""
        src = np.ascontiguousarray(sources, dtype=np.int64)
        k = len(src)
        r = np.zeros(max(k, 1), dtype=np.uint64)
        t, h, lg = (np.zeros(max(k, 1), dtype=np.float64) for _ in range(3))
        err = ctypes.create_string_buffer(ERR_LEN)
        kind = KIND[weights_as] if isinstance(weights_as, str) else int(weights_as)
        if self.L.orc_distance_profiles(self.h, int(weighted), kind, _p(src, c_int64), k,
                                        int(threads), _p(r, c_uint64), _p(t, c_double),
                                        _p(h, c_double), _p(lg, c_double), err, ERR_LEN):
            raise OracleError(err.value.decode())
        return dict(reached=r[:k], total=t[:k], harmonic=h[:k], longest=lg[:k])

    def single_source(self, s, weighted=False, weights_as="distance"):
        """(kind, dist_u64, dist_f64, sigma, delta, reached)."""
        n = self.n
        du = np.zeros(n, dtype=np.uint64)
        df = np.zeros(n, dtype=np.float64)
        sig = np.zeros(n, dtype=np.float64)
        dl = np.zeros(n, dtype=np.float64)
        reached = c_int64()
        kind = c_int32()
        err = ctypes.create_string_buffer(ERR_LEN)
        if self.L.orc_single_source(self.h, int(weighted), KIND[weights_as], int(s),
                                    ctypes.byref(reached), ctypes.byref(kind), _p(du, c_uint64),
                                    _p(df, c_double), _p(sig, c_double), _p(dl, c_double), err,
                                    ERR_LEN):
            raise OracleError(err.value.decode())
        return kind.value, du, df, sig, dl, reached.value


def ordered_sum_bits(items: int, threads: int) -> int:
    x = lib().orc_ordered_sum_test(items, threads)
    return int(np.float64(x).view(np.uint64))


def threads_for(requested: int, items: int, work: int) -> int:
    return int(lib().orc_threads_for(requested, items, work))


def available_cores() -> int:
    return int(lib().orc_available_cores())
