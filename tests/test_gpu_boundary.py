"""The C boundary on the device (GPU).

- A C99 program linked against libnetbc.so and run, as the reference's
  tools/check_c_abi.sh runs crates/capi/tests/c/smoke.c:292-312: the
  path 0-1-2-3 gives bc[1] = bc[2] = 2, a bad node index comes back as a
  non-zero status with the reference's message, net_graph_free(NULL) is a
  no-op, and a weighted call on an unweighted graph fails with its message.
- Device failures come back as status 1 with a message and leave the
  handle usable (crates/capi/src/lib.rs:73-85 turns panics into messages
  the same way): device allocations are refused through the library's
  failure-injection knob NETBC_FAIL_ALLOC_BYTES, before hipMalloc runs.
- netbc_calibration_copy refuses host and short buffers with a message.
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from knappnet_amd import _native, generate as gen
from knappnet_amd.engine import EngineError, Graph
from oracle.oracle import OracleGraph

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent

SMOKE = r"""
#include <stdio.h>
#include <string.h>
#include "netbc.h"

int main(void) {
    int64_t src[] = {0, 1, 2};
    int64_t tgt[] = {1, 2, 3};
    char err[256];
    GraphHandle *g = NULL;
    if (net_graph_new(4, src, tgt, NULL, 3, 0, &g, err, sizeof err) != 0) {
        fprintf(stderr, "graph_new failed: %s\n", err);
        return 1;
    }
    double bc[4];
    if (net_betweenness(g, 0, 0, 0, 1, bc, err, sizeof err) != 0) {
        fprintf(stderr, "betweenness failed: %s\n", err);
        return 2;
    }
    printf("bc %g %g %g %g\n", bc[0], bc[1], bc[2], bc[3]);
    if (bc[1] != 2.0 || bc[2] != 2.0 || bc[0] != 0.0 || bc[3] != 0.0) return 3;
    /* errors come back as messages, not crashes */
    int64_t bad[] = {9};
    GraphHandle *h = NULL;
    if (net_graph_new(2, bad, bad, NULL, 1, 1, &h, err, sizeof err) == 0) return 4;
    printf("bad index: %s\n", err);
    if (h != NULL) return 5;
    if (net_betweenness(g, 1, 0, 0, 1, bc, err, sizeof err) == 0) return 6;
    printf("weighted: %s\n", err);
    if (net_graph_node_count(g) != 4 || net_graph_edge_count(g) != 3) return 7;
    net_graph_free(g);
    net_graph_free(NULL);
    printf("ok\n");
    return 0;
}
"""


def test_c_program_linked_and_run(tmp_path):
    src = tmp_path / "smoke.c"
    src.write_text(SMOKE)
    exe = tmp_path / "smoke"
    lib_dir = _native.DEVICE_LIB.parent
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I",
                    str(ROOT / "include"), str(src), "-L", str(lib_dir), "-lnetbc",
                    f"-Wl,-rpath,{lib_dir}", "-o", str(exe)], check=True)
    # -lnetbc links libnetbc.so from the package folder
    ldd = subprocess.run(["ldd", str(exe)], capture_output=True, text=True, check=True).stdout
    assert str(lib_dir / "libnetbc.so") in ldd
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(run.stdout, run.stderr, flush=True)
    assert run.returncode == 0, run.stderr
    lines = run.stdout.splitlines()
    assert lines[0] == "bc 0 2 2 0"
    assert lines[1] == "bad index: edge 0 refers to node index 9, but valid indices are 0..2"
    assert lines[2] == "weighted: the graph has no edge weights"
    assert lines[3] == "ok"


@pytest.fixture
def refuse_alloc():
    def set_limit(b):
        if b is None:
            os.environ.pop("NETBC_FAIL_ALLOC_BYTES", None)
        else:
            os.environ["NETBC_FAIL_ALLOC_BYTES"] = str(int(b))
    yield set_limit
    os.environ.pop("NETBC_FAIL_ALLOC_BYTES", None)


def test_device_out_of_memory_is_a_message_and_the_handle_recovers(refuse_alloc):
    el = gen.rmat(12, 16, directed=True, seed=3)
    want = OracleGraph(el.n, el.src, el.tgt, None, True).betweenness(threads=0)
    # first call on a fresh handle: the device graph build itself fails
    g = Graph(el.n, el.src, el.tgt, None, True)
    refuse_alloc(1)
    with pytest.raises(EngineError, match=r"^device error: out of device memory allocating "
                                          r"\d+ bytes$"):
        g.betweenness()
    refuse_alloc(None)
    got, _ = g.betweenness_sources(exact=True)
    assert np.array_equal(got, want)
    # a second handle: the graph build (buffers well below 4 MiB) succeeds,
    # the engine workspace (4 slots of n x 64 doubles: 8 MiB rows) is
    # refused; with the knob cleared the same handle runs bit-exact
    g2 = Graph(el.n, el.src, el.tgt, None, True)
    refuse_alloc(4 << 20)
    with pytest.raises(EngineError, match=r"^device error: out of device memory allocating "
                                          r"8388608 bytes$"):
        g2.betweenness_sources(exact=True, slots=4)
    refuse_alloc(None)
    again, _ = g2.betweenness_sources(exact=True, slots=4)
    assert np.array_equal(again, want)
    # a batch of graphs reports the failure the same way
    from knappnet_amd.engine import betweenness_batch
    refuse_alloc(1)
    with pytest.raises(EngineError, match=r"out of device memory"):
        betweenness_batch([(4, [0, 1, 2], [1, 2, 3])])
    refuse_alloc(None)
    parts, _ = betweenness_batch([(4, [0, 1, 2], [1, 2, 3])])
    assert list(parts[0]) == [0.0, 2.0, 2.0, 0.0]


def test_calibration_copy_checks_its_buffers():
    lib = _native.device()
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(1 << 20)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(1 << 20)) == 0
    err = ctypes.create_string_buffer(256)
    try:
        assert lib.netbc_calibration_copy(b, a, 1 << 20, err, 256) == 0
        host = ctypes.create_string_buffer(1 << 12)
        assert lib.netbc_calibration_copy(b, ctypes.cast(host, ctypes.c_void_p), 1 << 12, err,
                                          256) == 1
        assert err.value == b"the source is not device memory"
        assert lib.netbc_calibration_copy(b, a, 2 << 20, err, 256) == 1
        assert err.value == b"the destination allocation is smaller than 2097152 bytes"
        # the library keeps working after the refused calls
        g = Graph(4, [0, 1, 2], [1, 2, 3], None, False)
        assert list(g.betweenness()) == [0.0, 2.0, 2.0, 0.0]
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def test_calibration_gather_checks_its_arguments():
    """The gather calibration reads one word per line of a device buffer:
    it refuses line sizes and buffer sizes it cannot cover and buffers that
    are not device memory of the stated size, and runs on a valid one."""
    lib = _native.device()
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(1 << 20)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(4 << 20)) == 0
    err = ctypes.create_string_buffer(256)
    try:
        assert lib.netbc_calibration_gather(b, a, 1 << 20, 128, err, 256) == 0, err.value
        assert lib.netbc_calibration_gather(b, a, 1 << 20, 96, err, 256) == 1
        assert err.value == b"line must be a power of two of at least 8 bytes"
        assert lib.netbc_calibration_gather(b, a, 3 << 18, 128, err, 256) == 1
        assert err.value == b"bytes / line must be a power of two"
        assert lib.netbc_calibration_gather(b, a, 2 << 20, 128, err, 256) == 1
        assert err.value == b"the source allocation is smaller than 2097152 bytes"
        assert lib.netbc_calibration_gather(a, b, 1 << 20, 128, err, 256) == 1
        assert err.value == b"the destination allocation is smaller than 4194304 bytes"
        g = Graph(4, [0, 1, 2], [1, 2, 3], None, False)
        assert list(g.betweenness()) == [0.0, 2.0, 2.0, 0.0]
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
