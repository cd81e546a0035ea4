// Internal declarations shared by the device library's translation units.
// Nothing here is part of the public C ABI (include/netbc.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace netbc {

constexpr int kLanes = 64;        // sources per batch = wavefront width on gfx950
constexpr int kBlock = 256;       // threads per block for the batched kernels
constexpr uint32_t kNone = 0xFFFFFFFFu;
#ifndef NETBC_HUB_IN
#define NETBC_HUB_IN 64
#endif
#ifndef NETBC_HUB_CHUNK
#define NETBC_HUB_CHUNK 512
#endif
constexpr uint32_t kHubIn = NETBC_HUB_IN;        // in-degree above which a node is a hub
constexpr uint32_t kHubChunk = NETBC_HUB_CHUNK;  // in-arcs per hub pull task
constexpr uint32_t kHeavyOut = 1024; // out-degree above which a row is split into tasks
constexpr uint32_t kChunkArcs = 256; // out-arcs per heavy-row task
constexpr int kLandmarks = 16;       // landmark searches for batching sources

// An error that reaches the C caller as status 1 and its message.
struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

#define NETBC_HIP(call)                                                              \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            throw ::netbc::Error(std::string("device error: ") + hipGetErrorString(e_) + \
                                 " (" #call ")");                                    \
    } while (0)

// Failure injection (tests): with NETBC_FAIL_ALLOC_BYTES=b in the
// environment, every device allocation of b bytes or more fails as if the
// device were out of memory, before hipMalloc is called.
inline bool alloc_refused(size_t bytes) {
    const char* s = std::getenv("NETBC_FAIL_ALLOC_BYTES");
    if (!s || !*s) return false;
    char* end = nullptr;
    const unsigned long long b = std::strtoull(s, &end, 10);
    return end != s && bytes >= b;
}

// A device allocation owned by host code.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Returns true when it made a new allocation (whose contents are
    // undefined, even where the address equals the old one's).
    bool alloc(size_t count) {
        if (count <= n && p) return false;
        release();
        size_t bytes = (count ? count : 1) * sizeof(T);
        hipError_t e = alloc_refused(bytes) ? hipErrorOutOfMemory : hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error("device error: out of device memory allocating " +
                        std::to_string(bytes) + " bytes");
        }
        n = count;
        return true;
    }
};

// Throws unless [p, p + bytes) lies inside one device allocation (buffers
// handed over by the caller; `what` names them in the message).
inline void check_device_range(const void* p, size_t bytes, const char* what) {
    hipPointerAttribute_t a{};
    if (!p || hipPointerGetAttributes(&a, p) != hipSuccess || a.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();  // a refused query must not surface in a later call
        throw Error(std::string("the ") + what + " is not device memory");
    }
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(std::string("the ") + what + " is not device memory");
    }
    const size_t used = (size_t)((const char*)p - (const char*)base);
    if (used > size || bytes > size - used)
        throw Error(std::string("the ") + what + " allocation is smaller than " +
                    std::to_string(bytes) + " bytes");
}

// Path-length classes, as the reference's Lengths enum (weights.rs:72-82).
enum LengthKind : int32_t { kUniform = 0, kInteger = 1, kReal = 2 };

// The device copy of a graph: out-CSR (rows sorted by target, parallel arcs
// in input order) and, for directed graphs, the in-CSR (CSC). For
// undirected graphs the in-CSR is the out-CSR.
struct DeviceGraph {
    uint64_t n = 0, m = 0, arcs = 0;
    bool directed = false;
    bool weighted = false;
    DevBuf<uint32_t> out_off, out_idx, in_off, in_idx;
    DevBuf<double> out_w, in_w;  // arc weights in CSR / CSC order
    DevBuf<uint32_t> arc_edge;   // input edge of every out-CSR arc
    // Hubs: nodes with more than kHubIn in-arcs. Their path counts (and,
    // bottom-up, their new lanes) are gathered in chunks of kHubChunk
    // in-arcs, one task per chunk and slot.
    DevBuf<uint8_t> hub_flag;          // per node
    DevBuf<uint32_t> hub_list;         // the hubs, ascending
    DevBuf<uint4> hub_task;            // per chunk: (hub, first in-arc, end in-arc, 0)
    uint32_t hubs = 0, hub_chunks = 0;
    uint32_t hub_in = 0, hub_chunk = 0;  // the thresholds the hub tables were cut with
    // Heavy rows: out-degree above kHeavyOut, expanded and swept in tasks
    // of kChunkArcs arcs.
    DevBuf<uint8_t> heavy_flag;        // per node
    uint32_t heavy = 0, max_chunks = 1;
    uint64_t max_out = 0;              // largest out-degree
    uint64_t max_in = 0;               // largest in-degree
    // Landmarks for batching sources: the nodes of largest in-degree
    // (ties by id), at most kLandmarks of them.
    std::vector<uint32_t> landmarks;
    // Hop distances from every node to the landmarks (a byte each, capped
    // at 15; n x kLandmarks), computed by the first run that batches
    // sources and kept: they depend on the graph alone.
    DevBuf<uint8_t> lm_hops;
    bool lm_hops_ready = false;
    const uint32_t* in_off_p() const { return directed ? in_off.p : out_off.p; }
    const uint32_t* in_idx_p() const { return directed ? in_idx.p : out_idx.p; }
    const double* in_w_p() const { return directed ? in_w.p : out_w.p; }
};

// Host graph handle: the validated edge list (input order) plus a lazily
// built device copy.
struct Graph {
    uint64_t n = 0, m = 0;
    bool directed = false;
    std::vector<uint32_t> src, tgt;
    std::vector<double> w;  // empty when unweighted
    bool has_weights = false;

    std::mutex mu;  // guards dev, the stream and the cached workspace
    bool dev_ready = false;  // device CSR and the batched engine's tables built
    bool dev_csr = false;    // device CSR built
    DeviceGraph dev;
    hipStream_t stream = nullptr;
    int device = -1;
    void* workspace = nullptr;  // engine scratch, freed with the handle
    void (*workspace_free)(void*) = nullptr;
    ~Graph();
};

// Length classification result (weights.rs:89-131), computed on device.
struct Lengths {
    LengthKind kind = kUniform;
    uint32_t max = 0;
    uint32_t min = 0;  // integer lengths: the shortest arc
    DevBuf<uint32_t> out_len, in_len;  // integer lengths, CSR / CSC order
    DevBuf<double> out_real, in_real;  // real lengths, CSR / CSC order
    double rmin = 0.0, rmax = 0.0;     // real lengths: shortest and longest arc
};

// Options and diagnostics of one betweenness run.
struct RunOptions {
    int precision = 0;      // 0: FP64 dependencies; 1: FP32 dependencies + Kahan
    int slots = 0;          // concurrent 64-source batches; 0 = auto
    int direction = 0;      // 0 auto, 1 top-down only, 2 bottom-up whenever possible
    double alpha = 0.0;     // bottom-up switch: frontier arcs > unvisited arcs / alpha (0 = default)
    int keep_sigma = 0;     // debug: keep per-source sigma and levels
    int exact = 0;          // 1: sum every row's successors in CSR order (no heavy-row split)
    int sweep_group = 0;    // entries per sweep task at most (0 = tuned default)
    int sweep_block = 0;    // threads per workgroup of the dependency sweep (0 = tuned)
    int step_block = 0;     // threads per workgroup of the bottom-up step and hub tasks
    int hub_in = 0;         // in-degree above which a node is a hub (0 = tuned)
    int hub_chunk = 0;      // in-arcs per hub task (0 = tuned)
    int bucket_width = 0;   // integer lengths: distances per bucket (0 = tuned)
};

struct RunDiag {
    double max_sigma = 0.0;
    int32_t sigma_overflow = 0;
    int32_t path_kind = 0;
    int64_t reached_nodes = 0;   // sum over sources of reached nodes
    int64_t reached_arcs = 0;    // sum over sources of out-arcs of reached nodes
    int64_t levels = 0;          // forward levels run (sum over super-batches)
    double kernel_ms = 0.0;      // device time of the search (events)
    double dominant_ms = 0.0;    // summed duration of the dominant kernel
    int64_t dominant_launches = 0;
    int64_t dominant_bytes = 0;  // algorithmic bytes of the dominant kernel, summed
    int32_t slots = 0;
    int64_t bottom_up_levels = 0;
    int64_t work[16] = {};  // engine work counters (diagnostic builds)
};

// Builds the device CSR/CSC of g (graph.hip). Caller holds g.mu.
void build_device_graph(Graph& g);
// The device CSR alone (the many-small-graphs engine reads nothing else).
void build_device_csr(Graph& g);
// Re-cuts the hub tables for other thresholds (hub_in: in-degree above
// which a node is a hub; hub_chunk: in-arcs per task); true if they changed.
bool ensure_hub_tables(Graph& g, uint32_t hub_in, uint32_t hub_chunk);

// The CSR/CSC of an unweighted graph of one fixed shape (n nodes, m edges,
// directed or not), rebuilt from new edges in place: fixed_csr_alloc
// allocates every buffer once, and fixed_csr_enqueue only launches kernels
// on the stream (no allocation, no memset, no wait on the host), so a
// stream capture can record it and a graph replay rebuild the arrays. The
// caller writes the edges (node ids, u32) into src/tgt first. The arrays
// equal build_device_graph's on rows [0, n); the target arrays hold `cap`
// slots, of which the rows use the first off[n].
struct FixedCsr {
    uint64_t n = 0, m = 0, cap = 0;  // cap: arc slots, m directed, 2m undirected
    bool directed = false;
    int bits = 0;                    // node-id bits of a key (ids 0..n are distinct)
    DevBuf<uint32_t> src, tgt;
    DevBuf<uint64_t> k0, k1;
    DevBuf<uint32_t> v0, v1, hist;
    DevBuf<uint32_t> scan[4];        // chunk sums of the scans, one buffer per level
    DevBuf<uint32_t> out_off, out_idx, in_off, in_idx;
    const uint32_t* in_off_p() const { return directed ? in_off.p : out_off.p; }
    const uint32_t* in_idx_p() const { return directed ? in_idx.p : out_idx.p; }
};
void fixed_csr_alloc(FixedCsr& c, uint64_t n, uint64_t m, bool directed);
void fixed_csr_enqueue(FixedCsr& c, hipStream_t st);

// Classifies lengths on the device, as path_lengths (weights.rs:89-131).
// Throws Error with the reference's message on bad weights.
void classify_lengths(Graph& g, bool weighted, int weights_as, Lengths& L);

// Runs for sources and targets (path_centrality.rs:421-593) on top of the
// batched engine: the dependency of a node counts 1 for itself only when it
// is a target, path endpoints can be counted, and real lengths can tie only
// when equal (Tie::Exact, paths/mod.rs:96-104).
struct SubsetRun {
    const uint8_t* d_target = nullptr;  // device, n flags; nullptr: every node
    bool endpoints = false;
    bool tie_exact = false;
    // edge values by NetworkX's subset edge split (path_centrality.rs:465-474)
    // instead of in proportion to path counts; node values of such a run are
    // not betweenness
    bool edge_split_nx = false;
};

// Runs the batched engine over item positions [0, k) with sources
// `sources` (nullptr: item i is source i) and writes unscaled sums into
// d_bc (device, n doubles) in the reference's group order. With d_eb
// (device, m doubles) it also writes the unscaled edge sums (reach.rs:
// 174-232): every arc's credit summed over the sources, then the arcs of
// each input edge added.
void run_batched(Graph& g, const Lengths& L, const int64_t* sources, int64_t k,
                 const RunOptions& opt, double* d_bc, RunDiag& diag, double* d_eb = nullptr,
                 const SubsetRun* sub = nullptr);

// Closeness or harmonic centrality of every node (closeness.rs:131-209)
// from the distance profiles of every node as a source (closeness.rs:
// 240-345), paths along the out-arcs of g (the caller orients g for the
// in and all modes). unreachable: 0 ignore, 1 Wasserman-Faust, 2 zero, 3
// the longest length plus one; unit: the common length of uniform weights
// (1 when unweighted), which hop counts are multiplied by.
struct ClosenessRun {
    bool harmonic = false;
    bool normalized = false;
    int unreachable = 0;
    double unit = 1.0;
    // Two-mode closeness instead (bipartite.rs:176-210): node types (device,
    // n flags, 1 = second mode) and the sizes of the two modes; unreachable
    // 0 ignore, 1 Wasserman-Faust.
    const uint8_t* d_types = nullptr;
    uint64_t first = 0, second = 0;
};

// Writes the n values to host `out`.
void run_closeness(Graph& g, const Lengths& L, const RunOptions& opt, const ClosenessRun& cr,
                   double* out, RunDiag& diag);

// Degree of every node (degree.rs:41-74; mode 1 out, 2 in, 3 all) into
// d_out (device, n doubles).
void degree_scores(Graph& g, int mode, bool loops, double* d_out);

// vals[v] /= (types[v] ? div1 : div0) where that divisor is positive, else
// NaN (device arrays, the graph's stream).
void divide_by_type(Graph& g, double* d_vals, const uint8_t* d_types, double div0, double div1);

// vals[v] *= flag[v] ? f_flagged : f_other on the graph's stream (flag:
// device, may be nullptr).
void scale_flagged(Graph& g, double* d_vals, size_t count, const uint8_t* d_flag,
                   double f_flagged, double f_other);

// Debug: per-source distances (hops or integer distances; UINT64_MAX when
// unreached), sigma and delta for k sources, each n long, host buffers.
void run_dump(Graph& g, const Lengths& L, const int64_t* sources, int64_t k,
              const RunOptions& opt, uint64_t* dist, double* sigma, double* delta);

// Scales d_bc in place as betweenness.rs:100-113 and copies it to host out.
void finish_scores(Graph& g, bool normalized, double* d_bc, double* out);

// Scales d_eb in place as reach.rs:207-224 and copies it to host out.
void finish_edge_scores(Graph& g, bool normalized, double* d_eb, double* out);

// Ensures g has a stream on the current device.
void ensure_stream(Graph& g);

// Stable LSD radix sort of (key, value) pairs on the low `bits` bits of the
// keys (graph.hip). Returns 0 if the result is in (k0, v0), 1 if in (k1, v1).
int sort_pairs_u64(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, size_t len, int bits,
                   hipStream_t st);

// A double as Rust's Display prints it (shortest round trip, positional),
// for error messages that quote a weight or a length.
std::string format_f64(double x);

// Bits needed for node ids 0..n-1 (at least 1).
int node_bits(uint64_t n);

}  // namespace netbc
