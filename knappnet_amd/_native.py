"""ctypes bindings of the in-tree native libraries.

``device()`` loads ``libnetbc.so`` (the HIP library, ``include/netbc.h``);
``generators()`` loads ``libnetgen.so``. A missing library raises
``RuntimeError`` naming the build command: there is no fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import (POINTER, Structure, c_char_p, c_double, c_int32, c_int64,
                    c_size_t, c_uint32, c_uint64, c_void_p)
from pathlib import Path

import os

PKG = Path(__file__).resolve().parent
# NETBC_LIB names another build of the device library (for A/B timing).
DEVICE_LIB = Path(os.environ.get("NETBC_LIB", PKG / "libnetbc.so"))
GEN_LIB = PKG / "libnetgen.so"

_device = None
_gen = None


class Options(Structure):
    """netbc_options (include/netbc.h)."""
    _fields_ = [("precision", c_int32), ("slots", c_int32), ("direction", c_int32),
                ("exact", c_int32), ("alpha", c_double), ("sweep_group", c_int32),
                ("sweep_block", c_int32), ("step_block", c_int32), ("hub_in", c_int32),
                ("hub_chunk", c_int32), ("bucket_width", c_int32), ("reserved", c_int32)]

    # the engine settings an autotuner may set (0 = the library's tuned value)
    TUNABLE = ("slots", "alpha", "sweep_group", "sweep_block", "step_block", "hub_in",
               "hub_chunk", "bucket_width")


class Diagnostics(Structure):
    """netbc_diagnostics (include/netbc.h)."""
    _fields_ = [("max_sigma", c_double), ("sigma_overflow", c_int32),
                ("path_kind", c_int32), ("reached_nodes", c_int64),
                ("reached_arcs", c_int64), ("levels", c_int64),
                ("kernel_ms", c_double), ("dominant_ms", c_double),
                ("dominant_launches", c_int64), ("slots", c_int32),
                ("bottom_up_levels", c_int32), ("build_ms", c_double), ("work", c_int64 * 16),
                ("dominant_bytes", c_int64)]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_
             if name != "work"}
        d["work"] = list(self.work)
        return d


def _load(path: Path, what: str) -> ctypes.CDLL:
    if not path.exists():
        raise RuntimeError(
            f"{what} is not built ({path} missing); run `python -m knappnet_amd.build`")
    return ctypes.CDLL(str(path))


def device() -> ctypes.CDLL:
    """The HIP betweenness library. Raises if it is not built."""
    global _device
    if _device is None:
        lib = _load(DEVICE_LIB, "the HIP device library")
        lib.net_graph_new.argtypes = [c_int64, POINTER(c_int64), POINTER(c_int64),
                                      POINTER(c_double), c_int64, c_int32,
                                      POINTER(c_void_p), c_char_p, c_size_t]
        lib.net_graph_new.restype = c_int32
        lib.net_graph_free.argtypes = [c_void_p]
        lib.net_graph_free.restype = None
        lib.net_graph_node_count.argtypes = [c_void_p]
        lib.net_graph_node_count.restype = c_int64
        lib.net_graph_edge_count.argtypes = [c_void_p]
        lib.net_graph_edge_count.restype = c_int64
        lib.net_betweenness.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_int32,
                                        POINTER(c_double), c_char_p, c_size_t]
        lib.net_betweenness.restype = c_int32
        lib.netbc_betweenness_sources.argtypes = [
            c_void_p, POINTER(c_int64), c_int64, c_int32, c_int32, c_int32,
            POINTER(Options), POINTER(c_double), POINTER(Diagnostics), c_char_p, c_size_t]
        lib.netbc_betweenness_sources.restype = c_int32
        lib.net_edge_betweenness.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_int32,
                                             POINTER(c_double), c_char_p, c_size_t]
        lib.net_edge_betweenness.restype = c_int32
        lib.net_betweenness_subset.argtypes = [
            c_void_p, POINTER(c_int64), c_int64, POINTER(c_int64), c_int64, c_int32, c_int32,
            c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, POINTER(c_double),
            POINTER(c_double), c_char_p, c_size_t]
        lib.net_betweenness_subset.restype = c_int32
        lib.netbc_betweenness_subset_ex.argtypes = [
            c_void_p, POINTER(c_int64), c_int64, POINTER(c_int64), c_int64, c_int32, c_int32,
            c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, POINTER(Options),
            POINTER(c_double), POINTER(c_double), POINTER(Diagnostics), c_char_p, c_size_t]
        lib.netbc_betweenness_subset_ex.restype = c_int32
        lib.net_centralization.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_int32,
                                           c_int32, c_int32, POINTER(c_double), c_char_p,
                                           c_size_t]
        lib.net_centralization.restype = c_int32
        lib.net_twomode_centrality.argtypes = [c_void_p, POINTER(c_int32), c_int32, c_int32,
                                               c_int32, c_int32, POINTER(c_double), c_char_p,
                                               c_size_t]
        lib.net_twomode_centrality.restype = c_int32
        lib.net_closeness.argtypes = [c_void_p, c_int32, c_int32, c_int32, c_int32, c_int32,
                                      c_int32, c_int32, POINTER(c_double), c_char_p, c_size_t]
        lib.net_closeness.restype = c_int32
        lib.netbc_closeness_ex.argtypes = [
            c_void_p, c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, POINTER(Options),
            POINTER(c_double), POINTER(Diagnostics), c_char_p, c_size_t]
        lib.netbc_closeness_ex.restype = c_int32
        lib.netbc_edge_betweenness_sources.argtypes = [
            c_void_p, POINTER(c_int64), c_int64, c_int32, c_int32, c_int32,
            POINTER(Options), POINTER(c_double), POINTER(Diagnostics), c_char_p, c_size_t]
        lib.netbc_edge_betweenness_sources.restype = c_int32
        lib.netbc_betweenness_batch.argtypes = [
            c_int64, POINTER(c_int64), POINTER(c_int64), POINTER(c_int64), POINTER(c_int64),
            POINTER(c_double), c_int32, c_int32, c_int32, c_int32, POINTER(Options),
            POINTER(c_double), POINTER(Diagnostics), c_char_p, c_size_t]
        lib.netbc_betweenness_batch.restype = c_int32
        lib.netbc_betweenness_batch_noalloc.argtypes = [
            c_int64, POINTER(c_int64), c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
            c_int64, c_void_p, c_void_p, c_char_p, c_size_t]
        lib.netbc_betweenness_batch_noalloc.restype = c_int32
        lib.netbc_batch_plan_new.argtypes = [
            c_int64, POINTER(c_int64), POINTER(c_int64), c_int32, c_int32, c_int32,
            POINTER(Options), POINTER(c_void_p), c_char_p, c_size_t]
        lib.netbc_batch_plan_new.restype = c_int32
        lib.netbc_batch_plan_run.argtypes = [
            c_void_p, POINTER(c_int64), POINTER(c_int64), POINTER(c_double),
            POINTER(Diagnostics), c_char_p, c_size_t]
        lib.netbc_batch_plan_run.restype = c_int32
        lib.netbc_batch_plan_free.argtypes = [c_void_p]
        lib.netbc_batch_plan_free.restype = None
        lib.netbc_state_dump.argtypes = [c_void_p, c_int32, c_int32, c_int32, POINTER(c_int64),
                                         c_int64, POINTER(c_uint64), POINTER(c_double),
                                         POINTER(c_double), c_char_p, c_size_t]
        lib.netbc_state_dump.restype = c_int32
        lib.netbc_graph_arc_count.argtypes = [c_void_p]
        lib.netbc_graph_arc_count.restype = c_int64
        lib.netbc_graph_csr.argtypes = [c_void_p, c_int32, POINTER(c_uint32), POINTER(c_uint32),
                                        POINTER(c_double), c_char_p, c_size_t]
        lib.netbc_graph_csr.restype = c_int32
        lib.netbc_calibration_copy.argtypes = [c_void_p, c_void_p, c_size_t, c_char_p, c_size_t]
        lib.netbc_calibration_copy.restype = c_int32
        lib.netbc_calibration_gather.argtypes = [c_void_p, c_void_p, c_size_t, c_size_t, c_char_p,
                                                 c_size_t]
        lib.netbc_calibration_gather.restype = c_int32
        lib.netbc_device_info.argtypes = [c_char_p, c_size_t]
        lib.netbc_device_info.restype = c_int32
        _device = lib
    return _device


def generators() -> ctypes.CDLL:
    """The host workload generators. Raises if they are not built."""
    global _gen
    if _gen is None:
        lib = _load(GEN_LIB, "the generator library")
        u32p = POINTER(c_uint32)
        lib.gen_philox4x32_10.argtypes = [u32p, u32p, u32p]
        lib.gen_philox4x32_10.restype = None
        lib.gen_stream_values.argtypes = [c_uint64, c_uint64, c_int32, c_uint64, c_int64,
                                          POINTER(c_uint64)]
        lib.gen_stream_values.restype = None
        lib.gen_permutation.argtypes = [c_uint64, c_uint64, c_int64, POINTER(c_int64)]
        lib.gen_permutation.restype = None
        lib.gen_gnm.argtypes = [c_int64, c_int64, c_int32, c_int32, c_uint64, c_uint64, u32p,
                                u32p]
        lib.gen_gnm.restype = c_int64
        lib.gen_small_world.argtypes = [c_int64, c_int64, c_double, c_uint64, c_uint64, u32p,
                                        u32p]
        lib.gen_small_world.restype = c_int64
        lib.gen_preferential.argtypes = [c_int64, c_int64, c_uint64, c_uint64, u32p, u32p]
        lib.gen_preferential.restype = c_int64
        lib.gen_rmat.argtypes = [c_int32, c_int64, c_double, c_double, c_double, c_int32,
                                 c_uint64, c_int32, POINTER(u32p), POINTER(u32p)]
        lib.gen_rmat.restype = c_int64
        lib.gen_free.argtypes = [c_void_p]
        lib.gen_free.restype = None
        lib.gen_fingerprint.argtypes = [u32p, u32p, c_int64]
        lib.gen_fingerprint.restype = c_uint64
        _gen = lib
    return _gen
