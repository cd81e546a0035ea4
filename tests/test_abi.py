"""The C ABI without a GPU: the library loads, exports every symbol that
include/netbc.h declares, the header compiles as C99 against it (the
reference's tools/check_c_abi.sh / smoke.c pattern), and the host-side error
paths return the reference's messages (lib.rs:60-164, error.rs:62-99)."""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from knappnet_amd import _native
from knappnet_amd.engine import EngineError, Graph

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "netbc.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(net\w+|netbc_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert "net_betweenness" in names and "net_graph_new" in names
    lib = ctypes.CDLL(str(_native.DEVICE_LIB))
    for name in names:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_native.DEVICE_LIB)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (\w+)$", out, flags=re.M))
    assert set(names) <= exported
    # nothing but the declared C entry points is exported with C names
    assert {s for s in exported if s.startswith("net")} == set(names)


def test_header_compiles_as_c99(tmp_path):
    src = tmp_path / "smoke.c"
    src.write_text(
        '#include "netbc.h"\n'
        "int main(void) {\n"
        "  GraphHandle *g = 0; char err[64]; double bc[1];\n"
        "  int64_t s[1] = {0}, t[1] = {0};\n"
        "  (void)net_graph_new(1, s, t, 0, 1, 0, &g, err, sizeof err);\n"
        "  (void)net_betweenness(g, 0, 0, 0, 0, bc, err, sizeof err);\n"
        "  net_graph_free(g); return 0;\n}\n")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", str(ROOT / "include"),
                    "-c", str(src), "-o", str(tmp_path / "smoke.o")], check=True)


def test_graph_handle_accessors():
    g = Graph(4, [0, 1, 2, 2], [1, 2, 3, 2], None, directed=False)
    assert g.node_count() == 4 and g.edge_count() == 4 and g.arc_count() == 7
    lib = _native.device()
    assert lib.net_graph_node_count(None) == -1
    assert lib.net_graph_edge_count(None) == -1
    lib.net_graph_free(None)


def test_construction_errors():
    with pytest.raises(EngineError, match=r"^edge 0 refers to node index 9, but valid "
                                          r"indices are 0..2$"):
        Graph(2, [9], [9], None, True)
    with pytest.raises(EngineError, match=r"^edge 1 refers to node index -1, but valid "
                                          r"indices are 0..3$"):
        Graph(3, [0, -1], [1, 2], None, True)
    with pytest.raises(EngineError, match=r"^edge 1 has a missing, NaN, or infinite weight$"):
        Graph(2, [0, 1], [1, 0], [1.0, np.inf], True)
    with pytest.raises(EngineError, match=r"^negative node or edge count$"):
        Graph(-1, [], [], None, True)


def test_call_errors_before_the_device():
    lib = _native.device()
    err = ctypes.create_string_buffer(128)
    out = (ctypes.c_double * 4)()
    assert lib.net_betweenness(None, 0, 0, 0, 0, out, err, 128) == 1
    assert err.value == b"graph handle is null"
    g = Graph(4, [0, 1, 2], [1, 2, 3], None, False)
    with pytest.raises(EngineError, match=r"^unknown weight kind code 7$"):
        g.betweenness(weights_as=7)
    with pytest.raises(EngineError, match=r"^the graph has no edge weights$"):
        g.betweenness(weighted=True)
    # err may be NULL or short: the message is truncated, never overflowed
    assert lib.net_betweenness(None, 0, 0, 0, 0, out, None, 0) == 1
    short = ctypes.create_string_buffer(b"\xff" * 8, 8)
    assert lib.net_betweenness(None, 0, 0, 0, 0, out, short, 8) == 1
    assert short.raw == b"graph h\x00"


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_no_device_is_an_error_not_a_fallback():
    g = Graph(4, [0, 1, 2], [1, 2, 3], None, False)
    with pytest.raises(EngineError, match="no HIP device"):
        g.betweenness()


def test_tuning_table_writer_compiles(tmp_path):
    """The autotuner's table is valid C++ with the layout the library reads."""
    import subprocess
    from knappnet_amd.autotune import write_table
    out = tmp_path / "tuning.h"
    s0 = dict(slots=64, alpha=1.0, sweep_group=32, sweep_block=256, step_block=512, hub_in=64,
              hub_chunk=512, bucket_width=1)
    write_table(out, [(0, 1, "R-MAT", s0),
                      (1, -1, "BA", dict(s0, slots=16, alpha=0.5, bucket_width=2)),
                      (2, 0, "graphs", dict(s0, sweep_block=1024))], "gfx950")
    src = tmp_path / "use.cpp"
    src.write_text('#include "tuning.h"\n'
                   'static_assert(sizeof(netbc::kTuning) / sizeof(netbc::kTuning[0]) == 3, "");\n'
                   'int main() { return netbc::kTuning[1].slots == 16 &&\n'
                   '    netbc::kTuning[1].bucket_width == 2 && netbc::kTuning[0].step_block == 512 &&\n'
                   '    netbc::kTuning[2].kind == 2 && netbc::kTuning[2].sweep_block == 1024 &&\n'
                   '    netbc::kTuning[2].slots == 0 ? 0 : 1; }\n')
    exe = tmp_path / "use"
    subprocess.run(["g++", "-std=c++17", "-I", str(tmp_path), str(src), "-o", str(exe)], check=True)
    assert subprocess.run([str(exe)]).returncode == 0
