// Many-small-graphs engine (small.hip): shared host/device descriptors.
#pragma once

#include <cstdint>
#include <vector>

#include "netbc_internal.h"

namespace netbc {

// Largest graph the one-workgroup engine takes (its frontier and visited
// bitmaps live in LDS: 3 x 8 bytes per node).
constexpr uint32_t kSmallMaxNodes = 2048;

// One graph of a batch, as laid out in the union graph.
struct SmallGraph {
    uint32_t n;        // nodes
    uint32_t node0;    // first node in the union graph
    uint32_t gs;       // source-group size ceil(n / 64) (parallel.rs:437)
    uint32_t groups;   // number of groups ceil(n / gs)
    uint64_t gsum_off; // offset of its group sums (groups x n doubles)
    double scale;      // betweenness.rs:100-113 scale
    uint32_t device;   // 1: computed by the one-workgroup engine
    uint32_t pad;
};

// One work item: whole groups [j0, j0 + ngroups) of graph g.
struct SmallItem {
    uint32_t g;
    uint32_t j0;
    uint32_t ngroups;
    uint32_t pad;
};

// Runs every graph with device == 1 in one persistent launch and writes
// the union graph's scores (scaled) to out_host. too_deep[g] is set for
// graphs whose search went past 254 levels (they must be recomputed).
void run_small_batch(Graph& g, const std::vector<SmallGraph>& graphs,
                     const std::vector<uint32_t>& node_graph, const RunOptions& opt,
                     double* out_host, std::vector<uint32_t>& too_deep, RunDiag& diag);

// Betweenness of many small unweighted graphs whose union is already on the
// device (out-CSR off/idx and, directed, in-CSR in_off/in_idx: `arcs`
// targets each, node ids consecutive across graphs), with no device
// allocation: scratch comes from a static device array (small.hip). Writes
// the scaled scores into device `out`. Bit-identical to run_small_batch.
void run_small_batch_noalloc(int64_t graphs, const int64_t* node_counts, bool directed,
                             bool normalized, const uint32_t* off, const uint32_t* idx,
                             const uint32_t* in_off, const uint32_t* in_idx, int64_t arcs,
                             double* out, hipStream_t st, RunDiag& diag);

}  // namespace netbc
