// Many small graphs in one persistent launch.
//
// Each work item is one batch of sources of one graph; a workgroup takes
// items from a queue (largest graphs first) and runs the whole batch by
// itself: the breadth-first levels (dense pulls over the graph, frontier
// bitmaps in LDS), the backward dependency sweep, and the group sums of
// the reference's ordered reduction (crates/engine/src/parallel.rs:383-503).
// Only workgroup barriers are needed, so a batch costs no kernel launches
// and no host round trips.
//
// Batches are cut along the reference's source groups: graph g with n
// nodes has groups of size gs = ceil(n / 64) (parallel.rs:437-440), and a
// batch holds floor(64 / gs) whole groups. Every group's sum
// acc = sum of delta_s(v) over its sources in order is therefore computed
// inside one workgroup, and a final pass adds the group sums of each node
// in group order and scales them (betweenness.rs:91-113). The arithmetic
// per source is the batched engine's, so results are bit-identical to a
// separate net_betweenness call per graph.

#include <algorithm>
#include <cstring>

#include "netbc_internal.h"
#include "small.h"

namespace netbc {
namespace {

#ifndef NETBC_SMALL_BLOCK
#define NETBC_SMALL_BLOCK 512
#endif
constexpr int kSmallBlock = NETBC_SMALL_BLOCK;
constexpr int kSmallWaves = kSmallBlock / 64;

__device__ __forceinline__ uint32_t srl32(uint32_t x, int j) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, j);
}
__device__ __forceinline__ uint64_t srl64(uint64_t x, int j) {
    uint32_t lo = srl32((uint32_t)x, j), hi = srl32((uint32_t)(x >> 32), j);
    return (uint64_t)hi << 32 | lo;
}

template <class C>
struct SDep;
template <>
struct SDep<double> {
    double s = 0.0;
    __device__ __forceinline__ void add(double x) { s += x; }
    __device__ __forceinline__ double value() const { return s; }
    __device__ __forceinline__ static double coef(double d, double sg) { return (1.0 + d) / sg; }
};
template <>
struct SDep<float> {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float x) {
        float y = x - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
    __device__ __forceinline__ static float coef(double d, double sg) {
        return (float)((1.0 + d) / sg);
    }
};

template <class C>
__global__ __launch_bounds__(kSmallBlock) void small_batch_kernel(
    const SmallItem* __restrict__ items, uint32_t nitems, uint32_t* __restrict__ counter,
    const SmallGraph* __restrict__ graphs, const uint32_t* __restrict__ off,
    const uint32_t* __restrict__ idx, const uint32_t* __restrict__ in_off,
    const uint32_t* __restrict__ in_idx, double* __restrict__ sig_all, C* __restrict__ coef_all,
    uint8_t* __restrict__ lvl_all, size_t nmax, double* __restrict__ gsum,
    uint32_t* __restrict__ too_deep, unsigned long long* __restrict__ stats) {
    __shared__ uint64_t vis[kSmallMaxNodes];
    __shared__ uint64_t fa[kSmallMaxNodes];
    __shared__ uint64_t fb[kSmallMaxNodes];
    __shared__ uint32_t item_s;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double* sig = sig_all + (size_t)blockIdx.x * nmax * kLanes;
    C* coef = coef_all + (size_t)blockIdx.x * nmax * kLanes;
    uint8_t* lvl = lvl_all + (size_t)blockIdx.x * nmax * kLanes;
    unsigned long long reached = 0, reached_arcs = 0;

    for (;;) {
        if (tid == 0) item_s = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t it = item_s;
        __syncthreads();
        if (it >= nitems) break;
        const SmallItem I = items[it];
        const SmallGraph G = graphs[I.g];
        const uint32_t n = G.n, base = G.node0;
        const uint32_t first = I.j0 * G.gs;
        const uint32_t nl = min(I.ngroups * G.gs, n - first);
        const uint64_t active = nl >= 64 ? ~0ull : ((1ull << nl) - 1);

        for (uint32_t v = tid; v < n; v += kSmallBlock) {
            vis[v] = 0;
            fa[v] = 0;
        }
        __syncthreads();
        if (tid < (int)nl) {
            uint32_t s = first + tid;
            atomicOr((unsigned long long*)&vis[s], 1ull << tid);
            atomicOr((unsigned long long*)&fa[s], 1ull << tid);
            sig[(size_t)s * kLanes + tid] = 1.0;
            lvl[(size_t)s * kLanes + tid] = 0;
        }
        __syncthreads();

        // Forward: level d -> d+1 by a pull over every node with unreached
        // lanes; sigma[w][l] sums sigma[u][l] over in-arcs u->w with u at
        // level d in lane l (in CSC order; exact integers).
        uint64_t* cur = fa;
        uint64_t* nxt = fb;
        int D = 0;
        bool deep = false;
        for (int d = 0;; d++) {
            int any = 0;
            for (uint32_t w = wave; w < n; w += kSmallWaves) {
                const uint64_t U = active & ~vis[w];
                uint64_t N = 0;
                if (U) {
                    double acc = 0.0;
                    const uint64_t a0 = in_off[base + w], a1 = in_off[base + w + 1];
                    for (uint64_t c = a0; c < a1; c += 64) {
                        const uint64_t a = c + lane;
                        const bool ok = a < a1;
                        const uint32_t u = ok ? in_idx[a] - base : 0;
                        const uint64_t mm = ok ? (cur[u] & U) : 0;
                        uint64_t bal = __ballot(mm != 0);
                        while (bal) {
                            const int j = __ffsll((unsigned long long)bal) - 1;
                            bal &= bal - 1;
                            const uint32_t uu = srl32(u, j);
                            const uint64_t m2 = srl64(mm, j);
                            N |= m2;
                            if ((m2 >> lane) & 1) acc += sig[(size_t)uu * kLanes + lane];
                        }
                    }
                    if ((N >> lane) & 1) {
                        sig[(size_t)w * kLanes + lane] = acc;
                        lvl[(size_t)w * kLanes + lane] = (uint8_t)(d + 1);
                    }
                }
                if (lane == 0) {
                    nxt[w] = N;
                    if (N) vis[w] |= N;
                }
                any |= N != 0;
            }
            if (!__syncthreads_or(any)) {
                D = d;
                break;
            }
            uint64_t* t = cur;
            cur = nxt;
            nxt = t;
            if (d + 2 >= 255) {
                deep = true;
                break;
            }
        }
        if (deep) {
            if (tid == 0) too_deep[I.g] = 1;
            __syncthreads();
            continue;
        }

        // Backward: levels D..1. fa = lanes at level d, fb = lanes at d+1.
        for (int d = D; d >= 1; d--) {
            for (uint32_t v = wave; v < n; v += kSmallWaves) {
                const uint8_t lv = lvl[(size_t)v * kLanes + lane];
                const uint64_t r = vis[v];
                const uint64_t Mv = __ballot(lv == d) & r;
                const uint64_t Sv = __ballot(lv == d + 1) & r;
                if (lane == 0) {
                    fa[v] = Mv;
                    fb[v] = Sv;
                }
            }
            __syncthreads();
            for (uint32_t v = wave; v < n; v += kSmallWaves) {
                const uint64_t M = fa[v];
                if (!M) continue;
                SDep<C> dv;
                if (d < D) {
                    const uint64_t a0 = off[base + v], a1 = off[base + v + 1];
                    for (uint64_t c = a0; c < a1; c += 64) {
                        const uint64_t a = c + lane;
                        const bool ok = a < a1;
                        const uint32_t x = ok ? idx[a] - base : 0;
                        const uint64_t mm = ok ? (fb[x] & M) : 0;
                        uint64_t bal = __ballot(mm != 0);
                        while (bal) {
                            const int j = __ffsll((unsigned long long)bal) - 1;
                            bal &= bal - 1;
                            const uint32_t xx = srl32(x, j);
                            const uint64_t m2 = srl64(mm, j);
                            if ((m2 >> lane) & 1) dv.add(coef[(size_t)xx * kLanes + lane]);
                        }
                    }
                }
                if ((M >> lane) & 1) {
                    const size_t r = (size_t)v * kLanes + lane;
                    const double s = sig[r];
                    const double delta = s * dv.value();
                    coef[r] = SDep<C>::coef(delta, s);
                    sig[r] = delta;
                }
            }
            __syncthreads();
        }

        // Group sums: for each whole group of the batch, in source order.
        for (uint32_t v = tid; v < n; v += kSmallBlock) {
            const uint64_t r = vis[v];
            reached += (unsigned long long)__popcll(r);
            reached_arcs += (unsigned long long)__popcll(r) * (off[base + v + 1] - off[base + v]);
            const double* row = sig + (size_t)v * kLanes;
            for (uint32_t j = 0; j < I.ngroups; j++) {
                const uint32_t l0 = j * G.gs;
                if (l0 >= nl) break;
                const uint32_t l1 = min(l0 + G.gs, nl);
                double acc = 0.0;
                for (uint32_t l = l0; l < l1; l++)
                    if (((r >> l) & 1) && first + l != v) acc += row[l];
                gsum[G.gsum_off + (size_t)(I.j0 + j) * n + v] = acc;
            }
        }
        __syncthreads();
    }
    // Reached nodes and arcs (diagnostics): summed over the wavefront, one
    // atomic per wavefront (one per thread made every thread of the launch
    // add into the same two addresses).
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        reached += (unsigned long long)__shfl_xor((long long)reached, o, 64);
        reached_arcs += (unsigned long long)__shfl_xor((long long)reached_arcs, o, 64);
    }
    if (lane == 0 && reached) atomicAdd(&stats[0], reached);
    if (lane == 0 && reached_arcs) atomicAdd(&stats[1], reached_arcs);
}

// bc[v] = sum over groups j (in order) of gsum[j][v], then the per-graph
// scale (betweenness.rs:91-113).
__global__ void small_reduce_kernel(const SmallGraph* __restrict__ graphs,
                                    const uint32_t* __restrict__ node_graph,
                                    const double* __restrict__ gsum, size_t total,
                                    double* __restrict__ out) {
    size_t u = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= total) return;
    const SmallGraph G = graphs[node_graph[u]];
    const uint32_t v = (uint32_t)(u - G.node0);
    double B = 0.0;
    for (uint32_t j = 0; j < G.groups; j++) B += gsum[G.gsum_off + (size_t)j * G.n + v];
    if (G.scale != 1.0) B *= G.scale;
    out[u] = B;
}

__global__ void zero_u32_kernel(uint32_t* p, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

}  // namespace

void run_small_batch(Graph& g, const std::vector<SmallGraph>& graphs,
                     const std::vector<uint32_t>& node_graph, const RunOptions& opt,
                     double* out_host, std::vector<uint32_t>& too_deep, RunDiag& diag) {
    hipStream_t st = g.stream;
    const DeviceGraph& d = g.dev;
    size_t total = g.n;
    // Work items, largest graphs first.
    std::vector<uint32_t> order(graphs.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (uint32_t)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return graphs[a].n > graphs[b].n; });
    std::vector<SmallItem> items;
    size_t gsum_total = 0;
    uint32_t nmax = 1;
    for (const SmallGraph& G : graphs) gsum_total = std::max<size_t>(gsum_total, G.gsum_off + (size_t)G.groups * G.n);
    for (uint32_t gi : order) {
        const SmallGraph& G = graphs[gi];
        if (G.n == 0 || !G.device) continue;
        nmax = std::max(nmax, G.n);
        uint32_t per = std::max<uint32_t>(1, 64 / G.gs);
        for (uint32_t j = 0; j < G.groups; j += per)
            items.push_back(SmallItem{gi, j, std::min(per, G.groups - j), 0});
    }
    DevBuf<SmallGraph> dgraphs;
    DevBuf<uint32_t> dnode_graph, dtoo_deep, counter;
    DevBuf<SmallItem> ditems;
    DevBuf<double> gsum, dout;
    DevBuf<unsigned long long> stats;
    dgraphs.alloc(graphs.size());
    dnode_graph.alloc(total);
    dtoo_deep.alloc(graphs.size());
    counter.alloc(1);
    ditems.alloc(items.size());
    gsum.alloc(gsum_total);
    dout.alloc(total);
    stats.alloc(2);
    NETBC_HIP(hipMemcpyAsync(dgraphs.p, graphs.data(), graphs.size() * sizeof(SmallGraph),
                             hipMemcpyHostToDevice, st));
    NETBC_HIP(hipMemcpyAsync(dnode_graph.p, node_graph.data(), total * 4, hipMemcpyHostToDevice,
                             st));
    if (!items.empty())
        NETBC_HIP(hipMemcpyAsync(ditems.p, items.data(), items.size() * sizeof(SmallItem),
                                 hipMemcpyHostToDevice, st));
    NETBC_HIP(hipMemsetAsync(dtoo_deep.p, 0, graphs.size() * 4, st));
    NETBC_HIP(hipMemsetAsync(stats.p, 0, 16, st));
    NETBC_HIP(hipMemsetAsync(gsum.p, 0, gsum_total * 8, st));

    int dev = 0;
    NETBC_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    NETBC_HIP(hipGetDeviceProperties(&prop, dev));
    int per_cu = 0;
    if (opt.precision)
        NETBC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, small_batch_kernel<float>, kSmallBlock, 0));
    else
        NETBC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, small_batch_kernel<double>, kSmallBlock, 0));
    per_cu = std::max(1, per_cu);
    size_t grid = std::min<size_t>((size_t)prop.multiProcessorCount * per_cu,
                                   std::max<size_t>(items.size(), 1));
    if (opt.slots > 0) grid = std::min<size_t>(grid, (size_t)opt.slots);
    DevBuf<double> sig;
    DevBuf<double> coef64;
    DevBuf<float> coef32;
    DevBuf<uint8_t> lvl;
    size_t per = (size_t)nmax * kLanes;
    sig.alloc(grid * per);
    lvl.alloc(grid * per);
    if (opt.precision) coef32.alloc(grid * per);
    else coef64.alloc(grid * per);

    hipEvent_t e0, e1;
    NETBC_HIP(hipEventCreate(&e0));
    NETBC_HIP(hipEventCreate(&e1));
    zero_u32_kernel<<<1, 64, 0, st>>>(counter.p, 1);
    NETBC_HIP(hipEventRecord(e0, st));
    if (!items.empty()) {
        if (opt.precision)
            small_batch_kernel<float><<<(unsigned)grid, kSmallBlock, 0, st>>>(
                ditems.p, (uint32_t)items.size(), counter.p, dgraphs.p, d.out_off.p,
                d.out_idx.p, d.in_off_p(), d.in_idx_p(), sig.p, coef32.p, lvl.p, nmax, gsum.p,
                dtoo_deep.p, stats.p);
        else
            small_batch_kernel<double><<<(unsigned)grid, kSmallBlock, 0, st>>>(
                ditems.p, (uint32_t)items.size(), counter.p, dgraphs.p, d.out_off.p,
                d.out_idx.p, d.in_off_p(), d.in_idx_p(), sig.p, coef64.p, lvl.p, nmax, gsum.p,
                dtoo_deep.p, stats.p);
        NETBC_HIP(hipGetLastError());
    }
    NETBC_HIP(hipEventRecord(e1, st));
    small_reduce_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(
        dgraphs.p, dnode_graph.p, gsum.p, total, dout.p);
    NETBC_HIP(hipGetLastError());
    NETBC_HIP(hipMemcpyAsync(out_host, dout.p, total * 8, hipMemcpyDeviceToHost, st));
    too_deep.assign(graphs.size(), 0);
    NETBC_HIP(hipMemcpyAsync(too_deep.data(), dtoo_deep.p, graphs.size() * 4,
                             hipMemcpyDeviceToHost, st));
    unsigned long long hs[2];
    NETBC_HIP(hipMemcpyAsync(hs, stats.p, 16, hipMemcpyDeviceToHost, st));
    NETBC_HIP(hipStreamSynchronize(st));
    float ms = 0.0f;
    NETBC_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    diag.kernel_ms += ms;
    diag.dominant_ms += ms;
    diag.dominant_launches += 1;
    diag.reached_nodes += (int64_t)hs[0];
    diag.reached_arcs += (int64_t)hs[1];
    diag.slots = (int32_t)grid;
}

}  // namespace netbc
