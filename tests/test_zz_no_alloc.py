"""netbc_betweenness_batch_noalloc against netbc_betweenness_batch (GPU).

The allocation-free entry takes the union graph already on the device (the
CSR net_graph_new builds, read back with netbc_graph_csr and uploaded here
with torch) and must return the batch entry's scores bit for bit. Two calls
in a row, the second with a smaller input (fewer, smaller graphs, directed
and normalized), check that the shared scratch starts each call clean.
Device memory is unchanged across calls once the library is loaded, and
the size limits and the device-side arc checks fail with messages.
"""
import ctypes

import numpy as np
import pytest

from knappnet_amd import _native, generate as gen
from knappnet_amd.engine import ERR_LEN, Graph, betweenness_batch

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def union(graphs, directed):
    """The union graph's CSR (and CSC when directed) as net_graph_new builds it."""
    ns = np.array([g[0] for g in graphs], dtype=np.int64)
    node0 = np.concatenate([[0], np.cumsum(ns)[:-1]])
    src = np.concatenate([np.asarray(g[1], np.int64) + node0[i] for i, g in enumerate(graphs)])
    tgt = np.concatenate([np.asarray(g[2], np.int64) + node0[i] for i, g in enumerate(graphs)])
    U = Graph(int(ns.sum()), src, tgt, None, directed)
    off, idx, _ = U.csr(False)
    ioff, iidx, _ = U.csr(True) if directed else (None, None, None)
    return ns, off, idx, ioff, iidx


def dev(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()


def noalloc(ns, off, idx, ioff, iidx, directed, normalized, out=None, arcs=None):
    lib = _native.device()
    total = int(ns.sum())
    d_off, d_idx, d_ioff, d_iidx = dev(off), dev(idx), dev(ioff), dev(iidx)
    if out is None:
        out = torch.full((max(total, 1),), -1.0, dtype=torch.float64, device="cuda")
    ptr = (lambda t: None if t is None else ctypes.c_void_p(t.data_ptr()))
    err = ctypes.create_string_buffer(ERR_LEN)
    ns = np.ascontiguousarray(ns, dtype=np.int64)
    rc = lib.netbc_betweenness_batch_noalloc(
        len(ns), ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(directed),
        int(normalized), ptr(d_off), ptr(d_idx), ptr(d_ioff), ptr(d_iidx),
        len(idx) if arcs is None else arcs, out if isinstance(out, ctypes.c_void_p) else
        ptr(out), None, err, ERR_LEN)
    if rc:
        return None, err.value.decode()
    return out.cpu().numpy()[:total], None


def free_device_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_noalloc_matches_batch_entry_on_two_calls():
    # call 1: 150 Watts-Strogatz graphs, n in [50, 2000], undirected
    els = gen.ws_batch(150, 23)
    graphs = [(e.n, e.src, e.tgt) for e in els]
    ns, off, idx, ioff, iidx = union(graphs, False)
    want = np.concatenate(betweenness_batch(graphs)[0])
    got, msg = noalloc(ns, off, idx, ioff, iidx, False, False)
    assert msg is None, msg
    assert np.array_equal(got, want), f"{int(np.sum(got != want))} scores differ"
    print("call 1:", len(graphs), "graphs,", int(ns.sum()), "nodes: bit-identical", flush=True)
    # call 2, smaller: 60 directed G(n, m) graphs, n in [1, 300], normalized
    rng = np.random.default_rng(9)
    graphs2 = []
    for i in range(60):
        n = int(rng.integers(1, 301))
        m = min(int(rng.integers(0, 4 * n + 1)), n * (n - 1))
        el = gen.gnm(n, m, directed=True, seed=100 + i)
        graphs2.append((n, el.src, el.tgt))
    ns2, off2, idx2, ioff2, iidx2 = union(graphs2, True)
    assert ns2.sum() < ns.sum() and len(graphs2) < len(graphs)
    want2 = np.concatenate(betweenness_batch(graphs2, directed=True, normalized=True)[0])
    before = free_device_bytes()
    got2, msg = noalloc(ns2, off2, idx2, ioff2, iidx2, True, True)
    assert msg is None, msg
    assert np.array_equal(got2, want2, equal_nan=True), \
        f"{int(np.sum(got2 != want2))} scores differ"
    print("call 2:", len(graphs2), "graphs,", int(ns2.sum()), "nodes: bit-identical", flush=True)
    # the second call ran on the loaded library: no device memory taken
    # (the inputs were uploaded before `before`)
    d = [dev(a) for a in (off2, idx2, ioff2, iidx2)]
    out = torch.empty(int(ns2.sum()), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    before = free_device_bytes()
    lib = _native.device()
    err = ctypes.create_string_buffer(ERR_LEN)
    ns2c = np.ascontiguousarray(ns2)
    rc = lib.netbc_betweenness_batch_noalloc(
        len(ns2c), ns2c.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 1, 1,
        *[ctypes.c_void_p(t.data_ptr()) for t in d], len(idx2), ctypes.c_void_p(out.data_ptr()),
        None, err, ERR_LEN)
    assert rc == 0, err.value
    assert free_device_bytes() == before
    assert np.array_equal(out.cpu().numpy(), want2, equal_nan=True)


def test_noalloc_limits_and_checks():
    graphs = [(4, [0, 1, 2], [1, 2, 3]), (3, [0, 1], [1, 2])]
    ns, off, idx, _, _ = union(graphs, False)
    got, msg = noalloc(ns, off, idx, None, None, False, False)
    assert msg is None and list(got) == [0.0, 2.0, 2.0, 0.0, 0.0, 1.0, 0.0]
    # sizes beyond the maximum
    big = np.array([2049], dtype=np.int64)
    _, msg = noalloc(big, np.zeros(2050, np.uint32), np.zeros(1, np.uint32), None, None, False,
                     False)
    assert msg == ("graph 0: 2049 nodes; the allocation-free entry takes graphs of at most "
                   "2048 nodes")
    many = np.ones(65537, dtype=np.int64)
    _, msg = noalloc(many, np.zeros(65538, np.uint32), np.zeros(1, np.uint32), None, None,
                     False, False)
    assert msg == "the allocation-free entry takes at most 65536 graphs per call"
    # a directed batch without its in-CSR
    _, msg = noalloc(ns, off, idx, None, None, True, False)
    assert msg == "a directed batch needs its in-CSR (in_offsets, in_targets)"
    # scores in host memory are refused before anything runs
    host = (ctypes.c_double * 8)()
    _, msg = noalloc(ns, off, idx, None, None, False, False,
                     out=ctypes.c_void_p(ctypes.addressof(host)))
    assert msg == "the score buffer is not device memory"
    # an arc that leaves its graph (target 6 lies in graph 1, listed in
    # graph 0's rows) is found by the kernel's checks
    bad = idx.copy()
    bad[0] = 6
    _, msg = noalloc(ns, off, bad, None, None, False, False)
    assert msg == ("graph 0: an offset runs backwards or past the arc count, or an arc leaves "
                   "the graph's nodes")
    # offsets past the arc count
    _, msg = noalloc(ns, off, idx, None, None, False, False, arcs=len(idx) - 2)
    assert msg is not None and msg.startswith("graph 1: an offset runs backwards")
    # and the next valid call is still right
    got, msg = noalloc(ns, off, idx, None, None, False, False)
    assert msg is None and list(got) == [0.0, 2.0, 2.0, 0.0, 0.0, 1.0, 0.0]
