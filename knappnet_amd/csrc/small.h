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

// The same engine split for repeated runs on one batch shape: small_setup
// allocates and uploads what the node counts fix (descriptors, work items,
// rows); small_enqueue launches the run on a union CSR (fill kernels, the
// persistent kernel, the ordered reduction: no allocation, no memset, no
// wait, so a stream capture can record it); small_fetch copies the scores,
// the too-deep flags and the reached counts back and waits.
struct SmallRun {
    size_t total = 0;       // union nodes
    size_t ngraphs = 0;
    uint32_t nitems = 0;
    uint32_t nmax = 1;
    size_t grid = 0;
    int precision = 0;
    int block = 512;        // threads per workgroup (256, 512 or 1024)
    DevBuf<SmallGraph> graphs;
    DevBuf<uint32_t> node_graph, too_deep, counter;
    DevBuf<SmallItem> items;
    DevBuf<double> gsum, out, sig, coef64;
    DevBuf<float> coef32;
    DevBuf<uint8_t> lvl;
    DevBuf<unsigned long long> stats;
};
void small_setup(SmallRun& r, const std::vector<SmallGraph>& graphs,
                 const std::vector<uint32_t>& node_graph, const RunOptions& opt, hipStream_t st);
void small_enqueue(SmallRun& r, const uint32_t* off, const uint32_t* idx, const uint32_t* in_off,
                   const uint32_t* in_idx, hipStream_t st, hipEvent_t e0 = nullptr,
                   hipEvent_t e1 = nullptr);
void small_fetch(SmallRun& r, double* out_host, std::vector<uint32_t>& too_deep,
                 unsigned long long stats[2], hipStream_t st);

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
