// Host-side tables of the device graph build (graph.hip): hub tasks cut
// from the in-rows and the heavy out-rows. Plain C++, no HIP, so the CPU
// tests compile it on its own.
//
// Offsets are u32: the reference accepts at most 2^31 - 1 edges
// (crates/engine/src/graph.rs:175-181), so an undirected graph stores at
// most 2^32 - 2 arcs and every offset fits. Sums of an offset and a step
// (a chunk end, a loop counter) do not: they are formed in 64 bits, so a
// row that ends within one step of 2^32 still ends.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace netbc {

// One hub task: in-arcs [first, end) of node `hub`.
struct HubTask {
    uint32_t hub, first, end;
};

// Nodes with more than `hub_in` in-arcs (flag[v] = 1, listed ascending in
// `hubs`) and their in-rows cut into tasks of `chunk` arcs, in node order.
// Returns the largest in-degree.
inline uint64_t cut_hub_tasks(const uint32_t* in_off, uint64_t n, uint32_t hub_in, uint32_t chunk,
                          std::vector<uint8_t>& flag, std::vector<uint32_t>& hubs,
                          std::vector<HubTask>& tasks) {
    flag.assign(std::max<uint64_t>(n, 1), 0);
    hubs.clear();
    tasks.clear();
    uint64_t maxdeg = 0;
    for (uint64_t v = 0; v < n; v++) {
        const uint64_t a0 = in_off[v], a1 = in_off[v + 1];
        maxdeg = std::max(maxdeg, a1 - a0);
        if (a1 - a0 <= hub_in) continue;
        flag[v] = 1;
        hubs.push_back((uint32_t)v);
        for (uint64_t a = a0; a < a1; a += chunk)
            tasks.push_back(HubTask{(uint32_t)v, (uint32_t)a, (uint32_t)std::min(a + chunk, a1)});
    }
    return maxdeg;
}

// Heavy out-rows (more than `heavy_out` arcs: flag[v] = 1); returns the
// largest out-degree. *chunks = chunk tasks of `chunk` arcs in the longest
// row (at least 1).
inline uint64_t scan_heavy_rows(const uint32_t* out_off, uint64_t n, uint32_t heavy_out,
                                uint32_t chunk, std::vector<uint8_t>& flag, uint32_t* heavy,
                                uint32_t* chunks) {
    flag.assign(std::max<uint64_t>(n, 1), 0);
    uint64_t maxdeg = 0;
    *heavy = 0;
    for (uint64_t v = 0; v < n; v++) {
        const uint64_t deg = (uint64_t)out_off[v + 1] - out_off[v];
        maxdeg = std::max(maxdeg, deg);
        if (deg > heavy_out) {
            flag[v] = 1;
            ++*heavy;
        }
    }
    *chunks = (uint32_t)std::max<uint64_t>(1, (maxdeg + chunk - 1) / chunk);
    return maxdeg;
}

}  // namespace netbc
