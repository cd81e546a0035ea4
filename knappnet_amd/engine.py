"""Python mirror of knappnet's ctypes harness over the C ABI.

``Graph`` follows the reference's ``validation/engine.py:92-128`` (the
harness its validation and benchmarks use): construct from an edge list,
call ``betweenness(weighted, normalized, threads, weights_as)``, get an
``EngineError`` carrying the library's message on failure. Every call goes
through ``libnetbc.so``; there is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

from . import _native

ERR_LEN = 1024
KIND = {"distance": 0, "strength": 1}


class EngineError(RuntimeError):
    """A non-zero status from the C ABI, with its message."""


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def _options(precision=0, slots=0, direction=0, exact=False, alpha=0.0, sweep_group=0,
             sweep_block=0, step_block=0, hub_in=0, hub_chunk=0, bucket_width=0):
    """netbc_options from keyword settings (0: the library's tuned value)."""
    return _native.Options(int(precision), int(slots), int(direction), int(bool(exact)),
                           float(alpha), int(sweep_group), int(sweep_block), int(step_block),
                           int(hub_in), int(hub_chunk), int(bucket_width), 0)


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(POINTER(ctype))


class Graph:
    """A graph handle of the device library (``net_graph_new``)."""

    def __init__(self, n: int, src, tgt, weights=None, directed: bool = False):
        self.lib = _native.device()
        self.n = int(n)
        self._src = _i64(src)
        self._tgt = _i64(tgt)
        m = len(self._src)
        self._w = None if weights is None else np.ascontiguousarray(
            np.asarray(weights, dtype=np.float64))
        self.handle = c_void_p()
        err = ctypes.create_string_buffer(ERR_LEN)
        status = self.lib.net_graph_new(
            self.n, _ptr(self._src, c_int64), _ptr(self._tgt, c_int64),
            None if self._w is None else _ptr(self._w, c_double), m, int(bool(directed)),
            ctypes.byref(self.handle), err, ERR_LEN)
        if status:
            self.handle = c_void_p()
            raise EngineError(err.value.decode())
        self.directed = bool(directed)

    def __del__(self):
        if getattr(self, "handle", None):
            self.lib.net_graph_free(self.handle)
            self.handle = c_void_p()

    def node_count(self) -> int:
        return int(self.lib.net_graph_node_count(self.handle))

    def edge_count(self) -> int:
        return int(self.lib.net_graph_edge_count(self.handle))

    def arc_count(self) -> int:
        return int(self.lib.netbc_graph_arc_count(self.handle))

    def betweenness(self, weighted: bool = False, normalized: bool = False, threads: int = 0,
                    weights_as: str = "distance") -> np.ndarray:
        """``net_betweenness``: scores of every node, in node order."""
        out = np.zeros(max(self.n, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        kind = KIND[weights_as] if isinstance(weights_as, str) else int(weights_as)
        if self.lib.net_betweenness(self.handle, int(bool(weighted)), kind,
                                    int(bool(normalized)), int(threads),
                                    _ptr(out, c_double), err, ERR_LEN):
            raise EngineError(err.value.decode())
        return out[: self.n]

    def edge_betweenness(self, weighted: bool = False, normalized: bool = False,
                         threads: int = 0, weights_as: str = "distance") -> np.ndarray:
        """``net_edge_betweenness``: scores of every edge, in input order."""
        m = self.edge_count()
        out = np.zeros(max(m, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        kind = KIND[weights_as] if isinstance(weights_as, str) else int(weights_as)
        if self.lib.net_edge_betweenness(self.handle, int(bool(weighted)), kind,
                                         int(bool(normalized)), int(threads),
                                         _ptr(out, c_double), err, ERR_LEN):
            raise EngineError(err.value.decode())
        return out[:m]

    def betweenness_subset(self, sources=None, targets=None, weighted: bool = False,
                           weights_as: str = "distance", endpoints: bool = False,
                           normalized: bool = False, rescale: int = 0, parallel: int = 0,
                           edge_split: int = 0, tie: int = 0, edges: bool = True,
                           options: dict | None = None):
        """``net_betweenness_subset``: (node values, edge values or None);
        with ``options`` (precision, slots, ...), ``netbc_betweenness_subset_ex``
        and a third item, the diagnostics. sources / targets None: every node."""
        def lst(x):
            if x is None:
                return None, -1, None
            a = _i64(x)
            return _ptr(a, c_int64), len(a), a
        sp, ns, _sa = lst(sources)
        tp, nt, _ta = lst(targets)
        m = self.edge_count()
        nodes = np.zeros(max(self.n, 1), dtype=np.float64)
        ev = np.zeros(max(m, 1), dtype=np.float64) if edges else None
        err = ctypes.create_string_buffer(ERR_LEN)
        kind = KIND[weights_as] if isinstance(weights_as, str) else int(weights_as)
        args = (self.handle, sp, ns, tp, nt, int(bool(weighted)), kind, int(bool(endpoints)),
                int(bool(normalized)), int(rescale), int(parallel), int(edge_split), int(tie))
        outs = (_ptr(nodes, c_double), None if ev is None else _ptr(ev, c_double))
        if options is None:
            if self.lib.net_betweenness_subset(*args, 0, *outs, err, ERR_LEN):
                raise EngineError(err.value.decode())
            return nodes[: self.n], None if ev is None else ev[:m]
        opt = _options(**options)
        diag = _native.Diagnostics()
        if self.lib.netbc_betweenness_subset_ex(*args, ctypes.byref(opt), *outs,
                                                ctypes.byref(diag), err, ERR_LEN):
            raise EngineError(err.value.decode())
        return nodes[: self.n], None if ev is None else ev[:m], diag.as_dict()

    def closeness(self, harmonic: bool = False, mode: int = 1, weighted: bool = False,
                  weights_as: str = "distance", normalized: bool = False, unreachable: int = 0,
                  options: dict | None = None):
        """``net_closeness``: closeness (or harmonic centrality) of every node.
        mode 1 out, 2 in, 3 all; unreachable 0 ignore, 1 Wasserman-Faust,
        2 zero, 3 the longest length plus one. With ``options``,
        ``netbc_closeness_ex``: returns (values, diagnostics dict)."""
        out = np.zeros(max(self.n, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        kind = KIND[weights_as] if isinstance(weights_as, str) else int(weights_as)
        args = (self.handle, int(bool(harmonic)), int(mode), int(bool(weighted)), kind,
                int(bool(normalized)), int(unreachable))
        if options is None:
            if self.lib.net_closeness(*args, 0, _ptr(out, c_double), err, ERR_LEN):
                raise EngineError(err.value.decode())
            return out[: self.n]
        opt = _options(**options)
        diag = _native.Diagnostics()
        if self.lib.netbc_closeness_ex(*args, ctypes.byref(opt), _ptr(out, c_double),
                                       ctypes.byref(diag), err, ERR_LEN):
            raise EngineError(err.value.decode())
        return out[: self.n], diag.as_dict()

    def centralization(self, measure: int = 0, mode: int = 3, loops: bool = True,
                       normalized: bool = True, convention: int = 0) -> float:
        """``net_centralization``: Freeman centralization. measure 0 degree, 1
        betweenness, 2 closeness; convention 0 igraph, 2 sna, 3 UCINET."""
        out = np.zeros(1, dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        if self.lib.net_centralization(self.handle, int(measure), int(mode), int(bool(loops)),
                                       int(bool(normalized)), int(convention), 0,
                                       _ptr(out, c_double), err, ERR_LEN):
            raise EngineError(err.value.decode())
        return float(out[0])

    def twomode_centrality(self, types, measure: int = 0, normalized: bool = True,
                           unreachable: int = 0):
        """``net_twomode_centrality``: measure 0 degree, 1 closeness, 2
        betweenness, normalized for two-mode data; types: 0 / 1 per node."""
        out = np.zeros(max(self.n, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        tp, _keep = None, None
        if types is not None:
            _keep = np.ascontiguousarray(types, dtype=np.int32)
            tp = _ptr(_keep, ctypes.c_int32)
        if self.lib.net_twomode_centrality(self.handle, tp, int(measure), int(bool(normalized)),
                                           int(unreachable), 0, _ptr(out, c_double), err,
                                           ERR_LEN):
            raise EngineError(err.value.decode())
        return out[: self.n]

    def _sources_call(self, fn, length, sources, weighted, normalized, weights_as, options):
        out = np.zeros(max(length, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        opt = _options(**options)
        diag = _native.Diagnostics()
        kind = KIND[weights_as] if isinstance(weights_as, str) else int(weights_as)
        if sources is None:
            src_p, k = None, self.n
        else:
            srcs = _i64(sources)
            src_p, k = _ptr(srcs, c_int64), len(srcs)
        if fn(self.handle, src_p, k, int(bool(weighted)), kind, int(bool(normalized)),
              ctypes.byref(opt), _ptr(out, c_double), ctypes.byref(diag), err, ERR_LEN):
            raise EngineError(err.value.decode())
        return out[:length], diag.as_dict()

    def betweenness_sources(self, sources=None, weighted: bool = False,
                            normalized: bool = False, weights_as: str = "distance", **options):
        """``netbc_betweenness_sources``: returns (scores, diagnostics dict).
        ``options``: the fields of netbc_options (precision, slots,
        direction, exact, alpha, sweep_group, sweep_block, step_block,
        hub_in, hub_chunk, bucket_width)."""
        return self._sources_call(self.lib.netbc_betweenness_sources, self.n, sources,
                                  weighted, normalized, weights_as, options)

    def edge_betweenness_sources(self, sources=None, weighted: bool = False,
                                 normalized: bool = False, weights_as: str = "distance",
                                 **options):
        """``netbc_edge_betweenness_sources``: returns (edge scores, diagnostics dict)."""
        return self._sources_call(self.lib.netbc_edge_betweenness_sources, self.edge_count(),
                                  sources, weighted, normalized, weights_as, options)

    def state_dump(self, sources, weighted: bool = False, weights_as: str = "distance",
                   precision: int = 0):
        """``netbc_state_dump``: (dist, sigma, delta), each k x n."""
        srcs = _i64(sources)
        k = len(srcs)
        dist = np.zeros((k, self.n), dtype=np.uint64)
        sigma = np.zeros((k, self.n), dtype=np.float64)
        delta = np.zeros((k, self.n), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        if self.lib.netbc_state_dump(self.handle, int(bool(weighted)), KIND[weights_as],
                                     int(precision), _ptr(srcs, c_int64), k,
                                     _ptr(dist, c_uint64), _ptr(sigma, c_double),
                                     _ptr(delta, c_double), err, ERR_LEN):
            raise EngineError(err.value.decode())
        return dist, sigma, delta

    def csr(self, incoming: bool = False):
        """The device CSR (or CSC): (offsets, targets, weights or None)."""
        arcs = self.arc_count()
        off = np.zeros(self.n + 1, dtype=np.uint32)
        tgt = np.zeros(max(arcs, 1), dtype=np.uint32)
        w = np.zeros(max(arcs, 1), dtype=np.float64) if self._w is not None else None
        err = ctypes.create_string_buffer(ERR_LEN)
        if self.lib.netbc_graph_csr(self.handle, int(bool(incoming)), _ptr(off, c_uint32),
                                    _ptr(tgt, c_uint32),
                                    None if w is None else _ptr(w, c_double), err, ERR_LEN):
            raise EngineError(err.value.decode())
        return off, tgt[:arcs], None if w is None else w[:arcs]


def _cat_i64(parts) -> np.ndarray:
    """The arrays concatenated as one contiguous int64 array (one
    conversion when they share an integer type)."""
    arrs = [np.asarray(p) for p in parts]
    if not arrs:
        return np.zeros(0, np.int64)
    if all(a.dtype == arrs[0].dtype for a in arrs) and arrs[0].dtype.kind in "iu":
        return np.ascontiguousarray(np.concatenate(arrs).astype(np.int64, copy=False))
    return _i64(np.concatenate([np.asarray(a, dtype=np.int64) for a in arrs]))


def betweenness_batch(graphs, directed: bool = False, weighted: bool = False,
                      normalized: bool = False, weights_as: str = "distance",
                      precision: int = 0, slots: int = 0, sweep_block: int = 0):
    """``netbc_betweenness_batch`` over a list of (n, src, tgt[, weights])
    graphs; returns (list of per-graph score arrays, diagnostics dict).
    sweep_block: the one-workgroup engine's workgroup size (256, 512 or
    1024; 0: the tuned size)."""
    lib = _native.device()
    ns = np.array([int(g[0]) for g in graphs], dtype=np.int64)
    ms = np.array([len(g[1]) for g in graphs], dtype=np.int64)
    src = _cat_i64([g[1] for g in graphs])
    tgt = _cat_i64([g[2] for g in graphs])
    w = None
    if weighted:
        w = np.ascontiguousarray(np.concatenate([np.asarray(g[3], dtype=np.float64)
                                                 for g in graphs]))
    out = np.zeros(max(int(ns.sum()), 1), dtype=np.float64)
    err = ctypes.create_string_buffer(ERR_LEN)
    opt = _options(precision=precision, slots=slots, sweep_block=sweep_block)
    diag = _native.Diagnostics()
    kind = KIND[weights_as] if isinstance(weights_as, str) else int(weights_as)
    if lib.netbc_betweenness_batch(len(graphs), _ptr(ns, c_int64), _ptr(ms, c_int64),
                                   _ptr(src, c_int64), _ptr(tgt, c_int64),
                                   None if w is None else _ptr(w, c_double), int(bool(directed)),
                                   int(bool(weighted)), kind, int(bool(normalized)),
                                   ctypes.byref(opt), _ptr(out, c_double), ctypes.byref(diag),
                                   err, ERR_LEN):
        raise EngineError(err.value.decode())
    parts = np.split(out[: int(ns.sum())], np.cumsum(ns)[:-1]) if len(graphs) else []
    return parts, diag.as_dict()


class BatchPlan:
    """``netbc_batch_plan_*``: repeated betweenness batches of one shape
    (graph count, node counts, edge counts), the device side recorded once
    into a HIP graph (``capture=True``) or launched kernel by kernel."""

    def __init__(self, node_counts, edge_counts, directed: bool = False,
                 normalized: bool = False, capture: bool = True, precision: int = 0,
                 slots: int = 0):
        self.lib = _native.device()
        self.ns = _i64(np.asarray(node_counts, dtype=np.int64))
        self.ms = _i64(np.asarray(edge_counts, dtype=np.int64))
        if len(self.ns) != len(self.ms):
            raise ValueError("node and edge counts differ in length")
        self.handle = ctypes.c_void_p()
        err = ctypes.create_string_buffer(ERR_LEN)
        opt = _options(precision=precision, slots=slots)
        if self.lib.netbc_batch_plan_new(len(self.ns), _ptr(self.ns, c_int64),
                                         _ptr(self.ms, c_int64), int(bool(directed)),
                                         int(bool(normalized)), int(bool(capture)),
                                         ctypes.byref(opt), ctypes.byref(self.handle), err,
                                         ERR_LEN):
            raise EngineError(err.value.decode())

    def run(self, graphs):
        """Scores of a list of (n, src, tgt) graphs of the plan's shape;
        returns (list of per-graph score arrays, diagnostics dict)."""
        if len(graphs) != len(self.ns):
            raise ValueError("graph count differs from the plan's")
        src = _i64(np.concatenate([np.asarray(g[1], dtype=np.int64) for g in graphs])
                   if graphs else np.zeros(0, np.int64))
        tgt = _i64(np.concatenate([np.asarray(g[2], dtype=np.int64) for g in graphs])
                   if graphs else np.zeros(0, np.int64))
        total = int(self.ns.sum())
        if len(src) != int(self.ms.sum()) or len(tgt) != len(src):
            raise ValueError("edge count differs from the plan's")
        out = np.zeros(max(total, 1), dtype=np.float64)
        err = ctypes.create_string_buffer(ERR_LEN)
        diag = _native.Diagnostics()
        if self.lib.netbc_batch_plan_run(self.handle, _ptr(src, c_int64), _ptr(tgt, c_int64),
                                         _ptr(out, c_double), ctypes.byref(diag), err, ERR_LEN):
            raise EngineError(err.value.decode())
        parts = np.split(out[:total], np.cumsum(self.ns)[:-1]) if len(self.ns) else []
        return parts, diag.as_dict()

    def close(self):
        if self.handle:
            self.lib.netbc_batch_plan_free(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_info() -> str:
    lib = _native.device()
    buf = ctypes.create_string_buffer(256)
    if lib.netbc_device_info(buf, 256):
        raise EngineError(buf.value.decode())
    return buf.value.decode()
