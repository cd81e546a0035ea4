// Device CSR/CSC construction and path-length classification.
//
// The reference builds its out-CSR with two stable counting sorts (first
// by target, then by source), so each row is sorted by target and parallel
// arcs keep input order (crates/engine/src/graph.rs:72-118); undirected
// edges become arcs (s,t),(t,s), a loop stays one arc (graph.rs:201-223);
// directed graphs also keep the in-CSR (graph.rs:197-200). Here the same
// order comes from one stable LSD radix sort of 64-bit keys
// (row << b | column) carrying the edge index, so the device arrays equal
// the reference's host arrays element for element.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "netbc_internal.h"
#include "row_tasks.h"

namespace netbc {
namespace {

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;
constexpr int kSortItems = 8;
constexpr int kSortTile = kBlock * kSortItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[kBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; i++) {
        if (i < wid) off += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

// Sum of each kScanTile chunk.
__global__ void scan_reduce_kernel(const uint32_t* __restrict__ in, size_t len,
                                   uint32_t* __restrict__ partial) {
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++)
        if (base + i < len) s += in[base + i];
    uint32_t tot;
    block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Exclusive scan of each chunk, offset by the scanned chunk sums.
__global__ void scan_apply_kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                  size_t len, const uint32_t* __restrict__ partial) {
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t vals[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        vals[i] = base + i < len ? in[base + i] : 0;
        s += vals[i];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, &tot) + (partial ? partial[blockIdx.x] : 0);
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        if (base + i < len) out[base + i] = run;
        run += vals[i];
    }
}

// In-place-safe exclusive scan of len u32 values (out may alias in).
void exclusive_scan(const uint32_t* in, uint32_t* out, size_t len, hipStream_t st) {
    if (len == 0) return;
    size_t blocks = (len + kScanTile - 1) / kScanTile;
    if (blocks == 1) {
        scan_apply_kernel<<<1, kBlock, 0, st>>>(in, out, len, nullptr);
        NETBC_HIP(hipGetLastError());
        return;
    }
    DevBuf<uint32_t> partial;
    partial.alloc(blocks);
    scan_reduce_kernel<<<blocks, kBlock, 0, st>>>(in, len, partial.p);
    NETBC_HIP(hipGetLastError());
    exclusive_scan(partial.p, partial.p, blocks, st);
    scan_apply_kernel<<<blocks, kBlock, 0, st>>>(in, out, len, partial.p);
    NETBC_HIP(hipGetLastError());
    NETBC_HIP(hipStreamSynchronize(st));  // partial is freed on return
}

// Digit histogram per tile, digit-major: hist[d * tiles + tile].
__global__ void radix_hist_kernel(const uint64_t* __restrict__ keys, size_t len, int shift,
                                  uint32_t* __restrict__ hist, uint32_t tiles) {
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    size_t base = (size_t)blockIdx.x * kSortTile;
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        size_t i = base + (size_t)r * kBlock + threadIdx.x;
        if (i < len) atomicAdd(&cnt[(keys[i] >> shift) & 255], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter of one tile: items are ranked in index order (round,
// wave, lane), so equal digits keep their relative order.
__global__ void radix_scatter_kernel(const uint64_t* __restrict__ kin,
                                     const uint32_t* __restrict__ vin,
                                     uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                     size_t len, int shift, const uint32_t* __restrict__ hist,
                                     uint32_t tiles) {
    __shared__ uint32_t base_d[256];
    __shared__ uint32_t wcnt[kBlock / 64][256];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    base_d[t] = hist[(size_t)t * tiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) wcnt[w][t] = 0;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kSortItems; r++) {
        size_t i = (size_t)blockIdx.x * kSortTile + (size_t)r * kBlock + t;
        bool valid = i < len;
        uint64_t k = valid ? kin[i] : 0;
        uint32_t v = valid ? vin[i] : 0;
        uint32_t d = (uint32_t)(k >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            bool set = (d >> bit) & 1u;
            uint64_t bb = __ballot(set);
            peers &= set ? bb : ~bb;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[wid][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = base_d[d] + rank;
            for (int w = 0; w < wid; w++) pos += wcnt[w][d];
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; w++) {
            add += wcnt[w][t];
            wcnt[w][t] = 0;
        }
        base_d[t] += add;
        __syncthreads();
    }
}

// Stable sort of (key, value) pairs on the low `bits` bits of the keys.
// Returns which buffer holds the result (0: k0/v0, 1: k1/v1).
int radix_sort_pairs(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, size_t len,
                     int bits, hipStream_t st) {
    if (len <= 1 || bits <= 0) return 0;
    uint32_t tiles = (uint32_t)((len + kSortTile - 1) / kSortTile);
    DevBuf<uint32_t> hist;
    hist.alloc((size_t)256 * tiles);
    int cur = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        uint64_t* ki = cur ? k1 : k0;
        uint32_t* vi = cur ? v1 : v0;
        uint64_t* ko = cur ? k0 : k1;
        uint32_t* vo = cur ? v0 : v1;
        radix_hist_kernel<<<tiles, kBlock, 0, st>>>(ki, len, shift, hist.p, tiles);
        NETBC_HIP(hipGetLastError());
        exclusive_scan(hist.p, hist.p, (size_t)256 * tiles, st);
        radix_scatter_kernel<<<tiles, kBlock, 0, st>>>(ki, vi, ko, vo, len, shift, hist.p, tiles);
        NETBC_HIP(hipGetLastError());
        cur ^= 1;
    }
    NETBC_HIP(hipStreamSynchronize(st));
    return cur;
}

__global__ void nonloop_flags_kernel(const uint32_t* __restrict__ s, const uint32_t* __restrict__ t,
                                     size_t m, uint32_t* __restrict__ flag) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < m) flag[e] = s[e] != t[e];
}

// Arc list of the reference's order before sorting: directed, one arc per
// edge; undirected, (s,t) then (t,s) unless a loop. Keys are
// (row << b | column); values the edge index.
__global__ void make_arcs_kernel(const uint32_t* __restrict__ s, const uint32_t* __restrict__ t,
                                 size_t m, int directed, int reverse, int b,
                                 const uint32_t* __restrict__ before,
                                 uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    uint64_t a = s[e], c = t[e];
    if (directed) {
        if (reverse) { uint64_t x = a; a = c; c = x; }
        key[e] = a << b | c;
        val[e] = (uint32_t)e;
        return;
    }
    size_t pos = e + before[e];
    key[pos] = a << b | c;
    val[pos] = (uint32_t)e;
    if (a != c) {
        key[pos + 1] = c << b | a;
        val[pos + 1] = (uint32_t)e;
    }
}

__global__ void row_count_kernel(const uint64_t* __restrict__ key, size_t arcs, int b,
                                 uint32_t* __restrict__ deg) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < arcs) atomicAdd(&deg[key[i] >> b], 1u);
}

__global__ void unpack_kernel(const uint64_t* __restrict__ key, const uint32_t* __restrict__ val,
                              size_t arcs, uint64_t mask, const double* __restrict__ w,
                              uint32_t* __restrict__ idx, double* __restrict__ wout) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= arcs) return;
    idx[i] = (uint32_t)(key[i] & mask);
    if (w) wout[i] = w[val[i]];
}

int bits_for(uint64_t n) {
    int b = 1;
    while (b < 64 && (1ull << b) < n) b++;
    return b;
}

inline unsigned grid_for(size_t items, int block = kBlock) {
    return (unsigned)((items + block - 1) / block);
}

void build_one(uint64_t n, uint64_t m, bool directed, bool reverse, const uint32_t* ds,
               const uint32_t* dt, const double* dw, const uint32_t* before, uint64_t arcs,
               DevBuf<uint32_t>& off, DevBuf<uint32_t>& idx, DevBuf<double>& wout,
               DevBuf<uint32_t>* arc_edge, hipStream_t st) {
    int b = bits_for(n);
    DevBuf<uint64_t> k0, k1;
    DevBuf<uint32_t> v0, v1;
    k0.alloc(arcs);
    k1.alloc(arcs);
    v0.alloc(arcs);
    v1.alloc(arcs);
    if (m) {
        make_arcs_kernel<<<grid_for(m), kBlock, 0, st>>>(ds, dt, m, directed, reverse, b, before,
                                                         k0.p, v0.p);
        NETBC_HIP(hipGetLastError());
    }
    int which = radix_sort_pairs(k0.p, v0.p, k1.p, v1.p, arcs, 2 * b, st);
    const uint64_t* key = which ? k1.p : k0.p;
    const uint32_t* val = which ? v1.p : v0.p;
    off.alloc(n + 1);
    NETBC_HIP(hipMemsetAsync(off.p, 0, (n + 1) * sizeof(uint32_t), st));
    if (arcs) {
        row_count_kernel<<<grid_for(arcs), kBlock, 0, st>>>(key, arcs, b, off.p);
        NETBC_HIP(hipGetLastError());
    }
    exclusive_scan(off.p, off.p, n + 1, st);
    idx.alloc(arcs);
    if (dw) wout.alloc(arcs);
    if (arcs) {
        unpack_kernel<<<grid_for(arcs), kBlock, 0, st>>>(key, val, arcs, (1ull << b) - 1, dw,
                                                         idx.p, dw ? wout.p : nullptr);
        NETBC_HIP(hipGetLastError());
    }
    if (arc_edge) {
        // the sorted values are the input edge of every arc (reach.rs:234-253)
        arc_edge->alloc(arcs);
        if (arcs)
            NETBC_HIP(hipMemcpyAsync(arc_edge->p, val, arcs * sizeof(uint32_t),
                                     hipMemcpyDeviceToDevice, st));
    }
    NETBC_HIP(hipStreamSynchronize(st));
}

// Classification flags (weights.rs:89-131).
struct ClassifyOut {
    unsigned long long first_bad;  // first edge (input order) with weight <= 0
    uint32_t not_uniform;
    uint32_t not_integer;
    uint32_t max;
    uint32_t min;  // shortest integer length
};

__global__ void first_nonpositive_kernel(const double* __restrict__ w, size_t m,
                                         ClassifyOut* out) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < m && w[e] <= 0.0) atomicMin(&out->first_bad, (unsigned long long)e);
}

__device__ __forceinline__ double arc_length(double w, int strength) {
    return strength ? 1.0 / w : w;
}

__global__ void classify_kernel(const double* __restrict__ w, size_t arcs, int strength,
                                ClassifyOut* out) {
    size_t a = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= arcs) return;
    double x = arc_length(w[a], strength);
    double x0 = arc_length(w[0], strength);
    if (!(x == x0)) out->not_uniform = 1;
    if (x - trunc(x) == 0.0 && x >= 1.0 && x <= 4294967295.0) {
        atomicMax(&out->max, (uint32_t)x);
        atomicMin(&out->min, (uint32_t)x);
    } else {
        out->not_integer = 1;
    }
}

__global__ void int_lengths_kernel(const double* __restrict__ w, size_t arcs, int strength,
                                   uint32_t* __restrict__ len) {
    size_t a = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < arcs) len[a] = (uint32_t)arc_length(w[a], strength);
}

// Real lengths in CSR (or CSC) order; with range, the shortest and longest
// as bit patterns (positive doubles order as their unsigned bits).
__global__ void real_lengths_kernel(const double* __restrict__ w, size_t arcs, int strength,
                                    double* __restrict__ len,
                                    unsigned long long* __restrict__ range) {
    size_t a = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= arcs) return;
    const double x = arc_length(w[a], strength);
    len[a] = x;
    if (range) {
        atomicMin(&range[0], (unsigned long long)__double_as_longlong(x));
        atomicMax(&range[1], (unsigned long long)__double_as_longlong(x));
    }
}

// Rust's Display for f64 (error.rs:81-85 prints the weight this way):
// shortest round-trip digits, positional notation.
std::string rust_f64(double x) {
    if (x != x) return "NaN";
    std::string sign = std::signbit(x) ? "-" : "";
    if (std::isinf(x)) return sign + "inf";
    x = std::fabs(x);
    if (x == 0.0) return sign + "0";
    char buf[64];
    for (int p = 1; p <= 17; p++) {
        snprintf(buf, sizeof buf, "%.*e", p - 1, x);
        if (strtod(buf, nullptr) == x) break;
    }
    std::string digits;
    const char* e = strchr(buf, 'e');
    for (const char* c = buf; c < e; c++)
        if (*c >= '0' && *c <= '9') digits.push_back(*c);
    int exp10 = atoi(e + 1);
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    std::string r;
    int nd = (int)digits.size();
    if (exp10 >= 0) {
        for (int i = 0; i <= exp10; i++) r.push_back(i < nd ? digits[i] : '0');
        if (nd > exp10 + 1) {
            r.push_back('.');
            r.append(digits, exp10 + 1, std::string::npos);
        }
    } else {
        r = "0.";
        r.append(-exp10 - 1, '0');
        r += digits;
    }
    return sign + r;
}

}  // namespace

std::string format_f64(double x) { return rust_f64(x); }

int sort_pairs_u64(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, size_t len, int bits,
                   hipStream_t st) {
    return radix_sort_pairs(k0, v0, k1, v1, len, bits, st);
}

int node_bits(uint64_t n) { return bits_for(n); }

Graph::~Graph() {
    if (workspace && workspace_free) workspace_free(workspace);
    if (stream) (void)hipStreamDestroy(stream);
}

void ensure_stream(Graph& g) {
    int dev = 0;
    NETBC_HIP(hipGetDevice(&dev));
    if (g.stream && g.device == dev) return;
    if (g.stream) {
        (void)hipStreamDestroy(g.stream);
        g.stream = nullptr;
    }
    NETBC_HIP(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    g.device = dev;
}

// Hub flags, list and tasks of g's device graph for the given thresholds
// (in_off: the in-row offsets on the host); sets max_in.
void cut_hub_tables(Graph& g, const uint32_t* ioff, uint32_t hub_in, uint32_t hub_chunk) {
    DeviceGraph& d = g.dev;
    std::vector<uint8_t> flag;
    std::vector<uint32_t> hub_list;
    std::vector<HubTask> cut;
    d.max_in = cut_hub_tasks(ioff, g.n, hub_in, hub_chunk, flag, hub_list, cut);
    std::vector<uint4> task(cut.size());
    for (size_t i = 0; i < cut.size(); i++)
        task[i] = make_uint4(cut[i].hub, cut[i].first, cut[i].end, 0);
    d.hubs = (uint32_t)hub_list.size();
    d.hub_list.alloc(hub_list.size());
    if (!hub_list.empty())
        NETBC_HIP(hipMemcpy(d.hub_list.p, hub_list.data(), hub_list.size() * 4,
                            hipMemcpyHostToDevice));
    d.hub_chunks = (uint32_t)task.size();
    d.hub_flag.alloc(flag.size());
    NETBC_HIP(hipMemcpy(d.hub_flag.p, flag.data(), flag.size(), hipMemcpyHostToDevice));
    d.hub_task.alloc(task.size());
    if (!task.empty())
        NETBC_HIP(hipMemcpy(d.hub_task.p, task.data(), task.size() * sizeof(uint4),
                            hipMemcpyHostToDevice));
    d.hub_in = hub_in;
    d.hub_chunk = hub_chunk;
}

bool ensure_hub_tables(Graph& g, uint32_t hub_in, uint32_t hub_chunk) {
    if (hub_in == 0 || hub_chunk < 64 || hub_chunk % 64)
        throw Error("hub thresholds out of range: in-degree " + std::to_string(hub_in) +
                    ", chunk " + std::to_string(hub_chunk));
    DeviceGraph& d = g.dev;
    if (d.hub_in == hub_in && d.hub_chunk == hub_chunk) return false;
    NETBC_HIP(hipStreamSynchronize(g.stream));  // the old tables may be in use
    std::vector<uint32_t> ioff(g.n + 1);
    NETBC_HIP(hipMemcpy(ioff.data(), d.in_off_p(), (g.n + 1) * 4, hipMemcpyDeviceToHost));
    cut_hub_tables(g, ioff.data(), hub_in, hub_chunk);
    return true;
}

void build_device_csr(Graph& g) {
    if (g.dev_csr) return;
    ensure_stream(g);
    hipStream_t st = g.stream;
    DeviceGraph& d = g.dev;
    d.n = g.n;
    d.m = g.m;
    d.directed = g.directed;
    d.weighted = g.has_weights;
    DevBuf<uint32_t> ds, dt, before;
    DevBuf<double> dw;
    ds.alloc(g.m);
    dt.alloc(g.m);
    if (g.m) {
        NETBC_HIP(hipMemcpyAsync(ds.p, g.src.data(), g.m * 4, hipMemcpyHostToDevice, st));
        NETBC_HIP(hipMemcpyAsync(dt.p, g.tgt.data(), g.m * 4, hipMemcpyHostToDevice, st));
    }
    if (g.has_weights) {
        dw.alloc(g.m);
        if (g.m)
            NETBC_HIP(hipMemcpyAsync(dw.p, g.w.data(), g.m * 8, hipMemcpyHostToDevice, st));
    }
    uint64_t arcs = g.m;
    if (!g.directed) {
        before.alloc(g.m + 1);
        if (g.m) {
            nonloop_flags_kernel<<<grid_for(g.m), kBlock, 0, st>>>(ds.p, dt.p, g.m, before.p);
            NETBC_HIP(hipGetLastError());
            exclusive_scan(before.p, before.p, g.m, st);
            uint32_t last = 0;
            NETBC_HIP(hipMemcpyAsync(&last, before.p + g.m - 1, 4, hipMemcpyDeviceToHost, st));
            NETBC_HIP(hipStreamSynchronize(st));
            uint64_t nonloops = (uint64_t)last + (g.src[g.m - 1] != g.tgt[g.m - 1]);
            arcs = g.m + nonloops;
        }
    }
    d.arcs = arcs;
    build_one(g.n, g.m, g.directed, false, ds.p, dt.p, g.has_weights ? dw.p : nullptr,
              before.p, arcs, d.out_off, d.out_idx, d.out_w, &d.arc_edge, st);
    if (g.directed)
        build_one(g.n, g.m, true, true, ds.p, dt.p, g.has_weights ? dw.p : nullptr, nullptr,
                  arcs, d.in_off, d.in_idx, d.in_w, nullptr, st);
    NETBC_HIP(hipStreamSynchronize(st));
    g.dev_csr = true;
}

void build_device_graph(Graph& g) {
    if (g.dev_ready) return;
    build_device_csr(g);
    DeviceGraph& d = g.dev;
    // Hub tables from the in-degrees.
    std::vector<uint32_t> ioff(g.n + 1);
    NETBC_HIP(hipMemcpy(ioff.data(), d.in_off_p(), (g.n + 1) * 4, hipMemcpyDeviceToHost));
    cut_hub_tables(g, ioff.data(), kHubIn, kHubChunk);
    // Heavy rows from the out-degrees.
    std::vector<uint32_t> ooff(g.n + 1);
    NETBC_HIP(hipMemcpy(ooff.data(), d.out_off.p, (g.n + 1) * 4, hipMemcpyDeviceToHost));
    std::vector<uint8_t> hflag;
    d.max_out = scan_heavy_rows(ooff.data(), g.n, kHeavyOut, kChunkArcs, hflag, &d.heavy,
                                &d.max_chunks);
    // Landmarks: largest in-degree first, ties by id.
    {
        std::vector<uint32_t> cand(g.n);
        for (uint64_t v = 0; v < g.n; v++) cand[v] = (uint32_t)v;
        const size_t want = std::min<size_t>(kLandmarks, g.n);
        auto by_in = [&](uint32_t a, uint32_t b) {
            uint32_t da = ioff[a + 1] - ioff[a], db = ioff[b + 1] - ioff[b];
            return da != db ? da > db : a < b;
        };
        std::partial_sort(cand.begin(), cand.begin() + want, cand.end(), by_in);
        d.landmarks.assign(cand.begin(), cand.begin() + want);
    }
    d.heavy_flag.alloc(hflag.size());
    NETBC_HIP(hipMemcpy(d.heavy_flag.p, hflag.data(), hflag.size(), hipMemcpyHostToDevice));
    g.dev_ready = true;
}

void classify_lengths(Graph& g, bool weighted, int weights_as, Lengths& L) {
    L.kind = kUniform;
    L.max = 0;
    if (!weighted) return;
    if (!g.has_weights) throw Error("the graph has no edge weights");
    hipStream_t st = g.stream;
    DeviceGraph& d = g.dev;
    DevBuf<ClassifyOut> res;
    res.alloc(1);
    ClassifyOut init{~0ull, 0, 0, 0, ~0u};
    NETBC_HIP(hipMemcpyAsync(res.p, &init, sizeof init, hipMemcpyHostToDevice, st));
    // The raw weights in input order are the out-CSR weights permuted; the
    // first offending edge index is found on the host copy's order by
    // scanning the device copy of the input-order weights.
    DevBuf<double> dw;
    dw.alloc(g.m);
    if (g.m) {
        NETBC_HIP(hipMemcpyAsync(dw.p, g.w.data(), g.m * 8, hipMemcpyHostToDevice, st));
        first_nonpositive_kernel<<<grid_for(g.m), kBlock, 0, st>>>(dw.p, g.m, res.p);
        NETBC_HIP(hipGetLastError());
    }
    int strength = weights_as == 1;
    if (d.arcs) {
        classify_kernel<<<grid_for(d.arcs), kBlock, 0, st>>>(d.out_w.p, d.arcs, strength, res.p);
        NETBC_HIP(hipGetLastError());
    }
    ClassifyOut out;
    NETBC_HIP(hipMemcpyAsync(&out, res.p, sizeof out, hipMemcpyDeviceToHost, st));
    NETBC_HIP(hipStreamSynchronize(st));
    if (out.first_bad != ~0ull) {
        uint64_t e = out.first_bad;
        throw Error("edge " + std::to_string(e) + " has weight " + rust_f64(g.w[e]) +
                    "; this algorithm treats weights as distances and requires them to be "
                    "strictly positive");
    }
    if (!out.not_uniform) return;  // Uniform -> breadth-first search
    if (!out.not_integer && out.max <= 1024u) {
        L.kind = kInteger;
        L.max = out.max;
        L.min = out.min;
        L.out_len.alloc(d.arcs);
        int_lengths_kernel<<<grid_for(d.arcs), kBlock, 0, st>>>(d.out_w.p, d.arcs, strength,
                                                                L.out_len.p);
        NETBC_HIP(hipGetLastError());
        if (d.directed) {
            L.in_len.alloc(d.arcs);
            int_lengths_kernel<<<grid_for(d.arcs), kBlock, 0, st>>>(d.in_w.p, d.arcs, strength,
                                                                    L.in_len.p);
            NETBC_HIP(hipGetLastError());
        }
    } else {
        L.kind = kReal;
        DevBuf<unsigned long long> range;
        range.alloc(2);
        const unsigned long long init_range[2] = {~0ull, 0ull};
        NETBC_HIP(hipMemcpyAsync(range.p, init_range, sizeof init_range, hipMemcpyHostToDevice,
                                 st));
        L.out_real.alloc(d.arcs);
        real_lengths_kernel<<<grid_for(d.arcs), kBlock, 0, st>>>(d.out_w.p, d.arcs, strength,
                                                                 L.out_real.p, range.p);
        NETBC_HIP(hipGetLastError());
        if (d.directed) {
            L.in_real.alloc(d.arcs);
            real_lengths_kernel<<<grid_for(d.arcs), kBlock, 0, st>>>(d.in_w.p, d.arcs, strength,
                                                                     L.in_real.p, nullptr);
            NETBC_HIP(hipGetLastError());
        }
        unsigned long long got[2];
        NETBC_HIP(hipMemcpyAsync(got, range.p, sizeof got, hipMemcpyDeviceToHost, st));
        NETBC_HIP(hipStreamSynchronize(st));
        memcpy(&L.rmin, &got[0], 8);
        memcpy(&L.rmax, &got[1], 8);
    }
    NETBC_HIP(hipStreamSynchronize(st));
}

// ---- fixed-shape build (FixedCsr) -------------------------------------------

namespace {

constexpr uint64_t kPadKey = ~0ull;  // an unused arc slot: sorts last, names no row

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, size_t count, uint32_t value) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

// Arc keys (row << b | column) at fixed slots: directed, slot e is edge e
// (reversed for the in-CSR); undirected, slots 2e and 2e + 1 are (s, t)
// and (t, s), or (s, s) and an unused slot for a loop. Slot order is the
// reference's arc order before sorting (graph.rs:201-223), so the stable
// sort keeps parallel arcs in input order. Values are the edge index.
__global__ void fixed_arcs_kernel(const uint32_t* __restrict__ s, const uint32_t* __restrict__ t,
                                  size_t m, int directed, int reverse, int b,
                                  uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    uint64_t a = s[e], c = t[e];
    if (directed) {
        if (reverse) {
            const uint64_t x = a;
            a = c;
            c = x;
        }
        key[e] = a << b | c;
        val[e] = (uint32_t)e;
        return;
    }
    key[2 * e] = a << b | c;
    val[2 * e] = (uint32_t)e;
    key[2 * e + 1] = a != c ? (c << b | a) : kPadKey;
    val[2 * e + 1] = (uint32_t)e;
}

// Arcs per row, unused slots (rows >= n) skipped.
__global__ void fixed_rows_kernel(const uint64_t* __restrict__ key, size_t slots, int b,
                                  uint64_t n, uint32_t* __restrict__ deg) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slots) return;
    const uint64_t row = key[i] >> b;
    if (row < n) atomicAdd(&deg[row], 1u);
}

__global__ void fixed_targets_kernel(const uint64_t* __restrict__ key, size_t slots,
                                     uint64_t mask, uint32_t* __restrict__ idx) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < slots) idx[i] = (uint32_t)(key[i] & mask);
}

// Exclusive scan with chunk sums in the preallocated buffers c.scan[level..]
// (no allocation, no wait).
void scan_fixed(const uint32_t* in, uint32_t* out, size_t len, FixedCsr& c, int level,
                hipStream_t st) {
    if (len == 0) return;
    const size_t blocks = (len + kScanTile - 1) / kScanTile;
    if (blocks == 1) {
        scan_apply_kernel<<<1, kBlock, 0, st>>>(in, out, len, nullptr);
        NETBC_HIP(hipGetLastError());
        return;
    }
    if (level >= 4 || c.scan[level].n < blocks) throw Error("fixed scan beyond its buffers");
    uint32_t* partial = c.scan[level].p;
    scan_reduce_kernel<<<blocks, kBlock, 0, st>>>(in, len, partial);
    NETBC_HIP(hipGetLastError());
    scan_fixed(partial, partial, blocks, c, level + 1, st);
    scan_apply_kernel<<<blocks, kBlock, 0, st>>>(in, out, len, partial);
    NETBC_HIP(hipGetLastError());
}

// Sorts c.cap (key, value) slots on the low 2 * c.bits key bits; returns
// which pair of buffers holds the result.
int sort_fixed(FixedCsr& c, hipStream_t st) {
    const size_t len = c.cap;
    const int bits = 2 * c.bits;
    if (len <= 1) return 0;
    const uint32_t tiles = (uint32_t)((len + kSortTile - 1) / kSortTile);
    int cur = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        uint64_t* ki = cur ? c.k1.p : c.k0.p;
        uint32_t* vi = cur ? c.v1.p : c.v0.p;
        uint64_t* ko = cur ? c.k0.p : c.k1.p;
        uint32_t* vo = cur ? c.v0.p : c.v1.p;
        radix_hist_kernel<<<tiles, kBlock, 0, st>>>(ki, len, shift, c.hist.p, tiles);
        NETBC_HIP(hipGetLastError());
        scan_fixed(c.hist.p, c.hist.p, (size_t)256 * tiles, c, 0, st);
        radix_scatter_kernel<<<tiles, kBlock, 0, st>>>(ki, vi, ko, vo, len, shift, c.hist.p,
                                                       tiles);
        NETBC_HIP(hipGetLastError());
        cur ^= 1;
    }
    return cur;
}

// One direction: keys, sort, row offsets (n + 1) and targets (cap slots).
void fixed_one(FixedCsr& c, bool reverse, uint32_t* off, uint32_t* idx, hipStream_t st) {
    const unsigned eg = grid_for(std::max<uint64_t>(c.m, 1));
    const unsigned sg = grid_for(std::max<uint64_t>(c.cap, 1));
    if (c.m) {
        fixed_arcs_kernel<<<eg, kBlock, 0, st>>>(c.src.p, c.tgt.p, c.m, c.directed ? 1 : 0,
                                                 reverse ? 1 : 0, c.bits, c.k0.p, c.v0.p);
        NETBC_HIP(hipGetLastError());
    }
    const int which = sort_fixed(c, st);
    const uint64_t* key = which ? c.k1.p : c.k0.p;
    fill_u32_kernel<<<grid_for(c.n + 1), kBlock, 0, st>>>(off, c.n + 1, 0u);
    NETBC_HIP(hipGetLastError());
    if (c.cap) {
        fixed_rows_kernel<<<sg, kBlock, 0, st>>>(key, c.cap, c.bits, c.n, off);
        NETBC_HIP(hipGetLastError());
        fixed_targets_kernel<<<sg, kBlock, 0, st>>>(key, c.cap, (1ull << c.bits) - 1, idx);
        NETBC_HIP(hipGetLastError());
    }
    scan_fixed(off, off, c.n + 1, c, 0, st);
}

}  // namespace

void fixed_csr_alloc(FixedCsr& c, uint64_t n, uint64_t m, bool directed) {
    if (n >= 0xFFFFFFFFull) throw Error("the batch exceeds 2^32 - 1 nodes in total");
    c.n = n;
    c.m = m;
    c.directed = directed;
    c.cap = directed ? m : 2 * m;
    if (c.cap > 0xFFFFFFFEull) throw Error("the batch exceeds 2^32 - 2 arcs in total");
    c.bits = bits_for(n + 1);  // ids 0..n, so the unused slots' row (all ones) is no node
    c.src.alloc(m);
    c.tgt.alloc(m);
    c.k0.alloc(c.cap);
    c.k1.alloc(c.cap);
    c.v0.alloc(c.cap);
    c.v1.alloc(c.cap);
    const size_t tiles = (c.cap + kSortTile - 1) / kSortTile;
    c.hist.alloc(256 * std::max<size_t>(tiles, 1));
    // chunk sums for the longest scan (the histogram or the offsets)
    size_t len = std::max<size_t>(256 * std::max<size_t>(tiles, 1), n + 1);
    for (int level = 0; level < 4; level++) {
        len = (len + kScanTile - 1) / kScanTile;
        c.scan[level].alloc(len);
        if (len <= 1) break;
    }
    c.out_off.alloc(n + 1);
    c.out_idx.alloc(c.cap);
    if (directed) {
        c.in_off.alloc(n + 1);
        c.in_idx.alloc(c.cap);
    }
}

void fixed_csr_enqueue(FixedCsr& c, hipStream_t st) {
    // every key slot is written by the arc kernel (an undirected loop's
    // unused slot gets the pad key), so nothing needs clearing
    fixed_one(c, false, c.out_off.p, c.out_idx.p, st);
    if (c.directed) fixed_one(c, true, c.in_off.p, c.in_idx.p, st);
}

}  // namespace netbc
