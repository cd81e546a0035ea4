"""Arc indices near 2^32 (CPU).

The reference accepts up to 2^31 - 1 edges (crates/engine/src/graph.rs:
175-181), so an undirected graph stores up to 2^32 - 2 arcs: every u32
offset is valid, but an offset plus a step is not. These tests pin that:

- the host tables of the device graph build (knappnet_amd/csrc/row_tasks.h,
  compiled here with g++ on its own) cut a row that ends at 2^32 - 2 into
  the right hub tasks and find the right heavy-row chunk count;
- no loop over arcs in the HIP sources steps a 32-bit counter (a step past
  2^32 would wrap and never end).
"""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "knappnet_amd" / "csrc"

DRIVER = r"""
#include <cstdio>
#include "row_tasks.h"
using namespace netbc;
int main() {
    // node 0: in-arcs [0, 10); node 1: [10, 2^32 - 600); node 2: the last
    // 598 arcs, ending at 2^32 - 2 (the largest arc count the reference accepts)
    const uint64_t top = 0xFFFFFFFEull;
    const uint32_t off[4] = {0, 10, (uint32_t)(top - 598), (uint32_t)top};
    std::vector<uint8_t> flag;
    std::vector<uint32_t> hubs;
    std::vector<HubTask> t;
    cut_hub_tasks(off, 3, 64, 512, flag, hubs, t);
    printf("hubs %zu flags %d %d %d\n", hubs.size(), flag[0], flag[1], flag[2]);
    printf("tasks %zu\n", t.size());
    const HubTask& a = t[t.size() - 2];
    const HubTask& b = t[t.size() - 1];
    printf("last %u %u %u | %u %u %u\n", a.hub, a.first, a.end, b.hub, b.first, b.end);
    uint32_t heavy = 0, chunks = 0;
    uint64_t mx = scan_heavy_rows(off, 3, 1024, 256, flag, &heavy, &chunks);
    printf("maxdeg %llu heavy %u chunks %u\n", (unsigned long long)mx, heavy, chunks);
    // one node holding every arc
    const uint32_t one[2] = {0, (uint32_t)top};
    mx = scan_heavy_rows(one, 1, 1024, 256, flag, &heavy, &chunks);
    printf("maxdeg %llu heavy %u chunks %u\n", (unsigned long long)mx, heavy, chunks);
    cut_hub_tasks(one, 1, 64, 512, flag, hubs, t);
    printf("tasks %zu last %u %u\n", t.size(), t.back().first, t.back().end);
    return 0;
}
"""


def test_row_tables_at_the_top_of_the_u32_range(tmp_path):
    src = tmp_path / "drv.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "drv"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I", str(CSRC), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True,
                         timeout=60).stdout.splitlines()
    top = 2**32 - 2
    mid = top - 598 - 10  # arcs of node 1
    assert out[0] == "hubs 2 flags 0 1 1"
    # node 1: ceil(mid / 512) tasks; node 2: 598 arcs -> 2 tasks (512 + 86)
    assert out[1] == f"tasks {-(-mid // 512) + 2}"
    assert out[2] == f"last 2 {top - 598} {top - 86} | 2 {top - 86} {top}"
    assert out[3] == f"maxdeg {mid} heavy 1 chunks {-(-mid // 256)}"
    assert out[4] == f"maxdeg {top} heavy 1 chunks {2**24}"
    assert out[5] == f"tasks {2**23} last {top - 510} {top}"


LOOP = re.compile(r"for\s*\(\s*(?:const\s+)?uint32_t\s+(\w+)\s*=[^;]*;[^;]*;\s*\1\s*\+=\s*"
                  r"(64|kHubChunk|kChunkArcs|chunk)\s*\)")


@pytest.mark.parametrize("name", ["batch.hip", "small.hip", "graph.hip", "capi.hip",
                                  "row_tasks.h"])
def test_no_32_bit_arc_loops(name):
    text = (CSRC / name).read_text()
    bad = []
    for mt in LOOP.finditer(text):
        line = text.count("\n", 0, mt.start()) + 1
        # loops over LDS staging buffers count entries, not arcs
        body = mt.group(0)
        if re.search(r"<\s*(cnt|tcnt)\b", body):
            continue
        bad.append(f"{name}:{line}: {body}")
    assert not bad, "32-bit counters stepping over arcs:\n" + "\n".join(bad)
