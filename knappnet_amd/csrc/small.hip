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
#include <mutex>
#include <type_traits>

#include "netbc_internal.h"
#include "small.h"
#include "tuning.h"

namespace netbc {
namespace {

#ifndef NETBC_SMALL_BLOCK
#define NETBC_SMALL_BLOCK 512
#endif
constexpr int kSmallBlock = NETBC_SMALL_BLOCK;

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

// Checked arc reads (graphs handed over as device arrays, which the host
// cannot validate): a row whose offsets run backwards or past the arc count
// is taken as empty, an arc leaving its graph's node range as absent, and
// the graph is reported in *bad (atomic minimum of graph indices).
struct ArcCheck {
    uint64_t arcs;
    uint32_t* bad;
};

// Row [a0, a1) of a node of graph g, empty (and g reported) when invalid.
template <bool kChecked>
__device__ __forceinline__ void row_bounds(const uint32_t* __restrict__ off, uint32_t node,
                                           uint32_t g, const ArcCheck& chk, uint64_t& a0,
                                           uint64_t& a1) {
    a0 = off[node];
    a1 = off[node + 1];
    if constexpr (kChecked) {
        if (a1 < a0 || a1 > chk.arcs) {
            atomicMin(chk.bad, g);
            a1 = a0;
        }
    }
}

// Node u - base of an arc's end, or n (no node) for an arc leaving the
// graph (reported when checked; a validated graph has none).
template <bool kChecked>
__device__ __forceinline__ uint32_t arc_end(uint32_t u, uint32_t base, uint32_t n, uint32_t g,
                                            const ArcCheck& chk) {
    const uint32_t x = u - base;
    if constexpr (kChecked) {
        if (x >= n) {
            atomicMin(chk.bad, g);
            return n;
        }
    }
    return x;
}

// Exclusive prefix of per-lane counts over the wavefront (total: their sum)
// and the lane whose items hold position t of the lanes' concatenation
// (the last lane c with pre[c] <= t). Idx: uint32_t where the rows of a
// graph are disjoint (their lengths sum to at most its 2^32 - 2 arcs),
// uint64_t for rows the host has not validated, which may overlap.
template <class Idx>
__device__ __forceinline__ Idx swave_prefix(Idx v, Idx& total) {
    const int lane = threadIdx.x & 63;
    Idx incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const Idx y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    total = __shfl(incl, 63, 64);
    return incl - v;
}
template <class Idx>
__device__ __forceinline__ int sowner_lane(Idx pre, Idx t) {
    int lo = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        const Idx pm = __shfl(pre, lo + step, 64);
        if (pm <= t) lo += step;
    }
    return lo;
}

// Rows loaded at once by the walks below.
constexpr int kSmallAhead = 8;

// One batch of sources (work item I of graph G) from start to finish in
// one workgroup: the breadth-first levels (pulls over the nodes with lanes
// still unreached, frontier bitmaps vis/fa/fb in LDS), then the backward
// dependency sweep. A wavefront takes 64 nodes at a time (lane c: node
// w0 + c) and walks their rows as one concatenation, 64 arcs per step, so
// that the memory accesses of 64 nodes overlap; hits are taken in arc
// order, so each node's sums run over its row in CSR order. Leaves
// delta_s(v) of lane l in sig[v * 64 + l] for the reached lanes of vis[v].
// Returns false (every thread) when the search passes 254 levels.
template <class C, bool kChecked, int kB = kSmallBlock>
__device__ bool run_small_item(const SmallItem& I, const SmallGraph& G,
                               const uint32_t* __restrict__ off, const uint32_t* __restrict__ idx,
                               const uint32_t* __restrict__ in_off,
                               const uint32_t* __restrict__ in_idx, double* __restrict__ sig,
                               C* __restrict__ coef, uint8_t* __restrict__ lvl, uint64_t* vis,
                               uint64_t* fa, uint64_t* fb, const ArcCheck& chk) {
    using Idx = std::conditional_t<kChecked, uint64_t, uint32_t>;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t n = G.n, base = G.node0;
    const uint32_t first = I.j0 * G.gs;
    const uint32_t nl = min(I.ngroups * G.gs, n - first);
    const uint64_t active = nl >= 64 ? ~0ull : ((1ull << nl) - 1);

    for (uint32_t v = tid; v < n; v += kB) {
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
    // level d in lane l (exact integers).
    uint64_t* cur = fa;
    uint64_t* nxt = fb;
    int D = 0;
    for (int d = 0;; d++) {
        int any = 0;
        for (uint32_t w0 = (uint32_t)wave * 64; w0 < n; w0 += kB) {
            const uint32_t w = w0 + (uint32_t)lane;
            const uint64_t U = w < n ? (active & ~vis[w]) : 0;
            uint64_t a0 = 0, a1 = 0;
            if (U) row_bounds<kChecked>(in_off, base + w, I.g, chk, a0, a1);
            Idx T;
            const Idx pre = swave_prefix((Idx)(a1 - a0), T);
            int curc = -1;  // node whose hits are being summed (wave-uniform)
            uint64_t ncur = 0, mine = 0;
            double acc = 0.0;
            auto close_pair = [&]() {
                if (curc < 0) return;
                const uint32_t xs = w0 + (uint32_t)curc;
                if ((ncur >> lane) & 1) {
                    sig[(size_t)xs * kLanes + lane] = acc;
                    lvl[(size_t)xs * kLanes + lane] = (uint8_t)(d + 1);
                }
                if (lane == curc) mine = ncur;
                acc = 0.0;
                ncur = 0;
            };
            for (Idx t0 = 0; t0 < T; t0 += 64) {
                const Idx t = t0 + (Idx)lane;
                const int c = sowner_lane(pre, t);
                const uint64_t ac = __shfl(a0, c, 64);
                const Idx pc = __shfl(pre, c, 64);
                const uint64_t Uc = __shfl(U, c, 64);
                uint32_t u = n;
                uint64_t m = 0;
                if (t < T) {
                    u = arc_end<kChecked>(in_idx[ac + (t - pc)], base, n, I.g, chk);
                    if (u < n) m = cur[u] & Uc;
                }
                uint64_t hits = __ballot(m != 0);
                while (hits) {
                    double val[kSmallAhead];
                    int hc[kSmallAhead];
                    uint64_t hm[kSmallAhead];
#pragma unroll
                    for (int k = 0; k < kSmallAhead; k++) {
                        val[k] = 0.0;
                        hc[k] = -1;
                        hm[k] = 0;
                        if (hits) {
                            const int j = __ffsll((unsigned long long)hits) - 1;
                            hits &= hits - 1;
                            const uint32_t uu = srl32(u, j);
                            hm[k] = srl64(m, j);
                            hc[k] = (int)srl32((uint32_t)c, j);
                            if ((hm[k] >> lane) & 1) val[k] = sig[(size_t)uu * kLanes + lane];
                        }
                    }
#pragma unroll
                    for (int k = 0; k < kSmallAhead; k++) {
                        if (hc[k] < 0) break;
                        if (hc[k] != curc) {
                            close_pair();
                            curc = hc[k];
                        }
                        ncur |= hm[k];
                        acc += val[k];
                    }
                }
            }
            close_pair();
            if (w < n) {
                nxt[w] = mine;
                if (mine) vis[w] |= mine;
            }
            any |= mine != 0;
        }
        if (!__syncthreads_or(any)) {
            D = d;
            break;
        }
        uint64_t* t = cur;
        cur = nxt;
        nxt = t;
        if (d + 2 >= 255) return false;
    }

    // Backward: levels D..1. fa = lanes at level d, fb = lanes at d+1.
    for (int d = D; d >= 1; d--) {
        // a wavefront reads 4 nodes' level rows at once (byte per lane)
        for (uint32_t v0 = (uint32_t)wave * 4; v0 < n; v0 += kB / 16) {
            uint8_t lv[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
                lv[q] = v0 + q < n ? lvl[(size_t)(v0 + q) * kLanes + lane] : (uint8_t)0xFF;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t v = v0 + (uint32_t)q;
                if (v >= n) break;
                const uint64_t r = vis[v];
                const uint64_t Mv = __ballot(lv[q] == d) & r;
                const uint64_t Sv = __ballot(lv[q] == d + 1) & r;
                if (lane == 0) {
                    fa[v] = Mv;
                    fb[v] = Sv;
                }
            }
        }
        __syncthreads();
        for (uint32_t v0 = (uint32_t)wave * 64; v0 < n; v0 += kB) {
            const uint32_t v = v0 + (uint32_t)lane;
            const uint64_t M = v < n ? fa[v] : 0;
            uint64_t a0 = 0, a1 = 0;
            if (M && d < D) row_bounds<kChecked>(off, base + v, I.g, chk, a0, a1);
            Idx T;
            const Idx pre = swave_prefix((Idx)(a1 - a0), T);
            bool done = M == 0;
            int cur_e = -1;  // node whose successors are being summed (wave-uniform)
            uint64_t Me = 0;
            double sv = 0.0;
            SDep<C> dv;
            auto finish = [&](uint32_t vv, uint64_t Mv, double s, double dvv) {
                if ((Mv >> lane) & 1) {
                    const size_t r = (size_t)vv * kLanes + lane;
                    const double delta = s * dvv;
                    coef[r] = SDep<C>::coef(delta, s);
                    sig[r] = delta;
                }
            };
            auto close_entry = [&]() {
                if (cur_e < 0) return;
                finish(v0 + (uint32_t)cur_e, Me, sv, dv.value());
                if (lane == cur_e) done = true;
                dv = SDep<C>();
            };
            auto open_entry = [&](int c) {
                cur_e = c;
                Me = srl64(M, c);
                sv = ((Me >> lane) & 1) ? sig[(size_t)(v0 + (uint32_t)c) * kLanes + lane] : 0.0;
            };
            for (Idx t0 = 0; t0 < T; t0 += 64) {
                const Idx t = t0 + (Idx)lane;
                const int c = sowner_lane(pre, t);
                const uint64_t ac = __shfl(a0, c, 64);
                const Idx pc = __shfl(pre, c, 64);
                const uint64_t Mc = __shfl(M, c, 64);
                uint32_t x = n;
                uint64_t mm = 0;
                if (t < T) {
                    x = arc_end<kChecked>(idx[ac + (t - pc)], base, n, I.g, chk);
                    if (x < n) mm = fb[x] & Mc;
                }
                uint64_t hits = __ballot(mm != 0);
                while (hits) {
                    C val[kSmallAhead];
                    int hc[kSmallAhead];
                    bool use[kSmallAhead];
#pragma unroll
                    for (int k = 0; k < kSmallAhead; k++) {
                        val[k] = 0;
                        hc[k] = -1;
                        use[k] = false;
                        if (hits) {
                            const int j = __ffsll((unsigned long long)hits) - 1;
                            hits &= hits - 1;
                            const uint32_t xx = srl32(x, j);
                            const uint64_t m2 = srl64(mm, j);
                            hc[k] = (int)srl32((uint32_t)c, j);
                            use[k] = (m2 >> lane) & 1;
                            if (use[k]) val[k] = coef[(size_t)xx * kLanes + lane];
                        }
                    }
#pragma unroll
                    for (int k = 0; k < kSmallAhead; k++) {
                        if (hc[k] < 0) break;
                        if (hc[k] != cur_e) {
                            close_entry();
                            open_entry(hc[k]);
                        }
                        if (use[k]) dv.add(val[k]);
                    }
                }
            }
            close_entry();
            // entries without successors in their lanes: delta = 0
            uint64_t rest = __ballot(!done);
            while (rest) {
                const int j = __ffsll((unsigned long long)rest) - 1;
                rest &= rest - 1;
                const uint64_t Mj = srl64(M, j);
                const uint32_t vj = v0 + (uint32_t)j;
                const double s = ((Mj >> lane) & 1) ? sig[(size_t)vj * kLanes + lane] : 0.0;
                finish(vj, Mj, s, SDep<C>().value());
            }
        }
        __syncthreads();
    }
    return true;
}

// The sum over sources s of group j of item I (in source order) of
// delta_s(v), v != s: the reference's per-group accumulator
// (parallel.rs:437-503) for node v. Also counts reached nodes and arcs.
__device__ __forceinline__ double group_sum(const SmallItem& I, const SmallGraph& G,
                                            const double* __restrict__ row, uint64_t r,
                                            uint32_t v, uint32_t j) {
    const uint32_t first = I.j0 * G.gs;
    const uint32_t nl = min(I.ngroups * G.gs, G.n - first);
    const uint32_t l0 = j * G.gs;
    const uint32_t l1 = min(l0 + G.gs, nl);
    double acc = 0.0;
    for (uint32_t l = l0; l < l1; l++)
        if (((r >> l) & 1) && first + l != v) acc += row[l];
    return acc;
}

// Wavefronts per SIMD the one-workgroup kernel's registers are capped for:
// as many as its 49 KB of LDS per workgroup lets a CU hold (3 workgroups
// of 256 or 512 threads, 1 of 1024). At 512 threads the cap (85 -> 80
// VGPRs) takes it from 2 to 3 workgroups per CU: 4,096 Watts-Strogatz
// graphs 0.651 -> 0.548 s.
constexpr int small_waves(int block) { return block == 256 ? 3 : block == 512 ? 6 : 4; }

template <class C, int kB>
__global__ __launch_bounds__(kB, small_waves(kB)) void small_batch_kernel(
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
    const int tid = threadIdx.x, lane = tid & 63;
    double* sig = sig_all + (size_t)blockIdx.x * nmax * kLanes;
    C* coef = coef_all + (size_t)blockIdx.x * nmax * kLanes;
    uint8_t* lvl = lvl_all + (size_t)blockIdx.x * nmax * kLanes;
    unsigned long long reached = 0, reached_arcs = 0;
    const ArcCheck chk{0, nullptr};

    for (;;) {
        if (tid == 0) item_s = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t it = item_s;
        __syncthreads();
        if (it >= nitems) break;
        const SmallItem I = items[it];
        const SmallGraph G = graphs[I.g];
        if (!run_small_item<C, false, kB>(I, G, off, idx, in_off, in_idx, sig, coef, lvl, vis, fa,
                                      fb, chk)) {
            if (tid == 0) too_deep[I.g] = 1;
            __syncthreads();
            continue;
        }
        // Group sums: for each whole group of the batch, in source order.
        const uint32_t first = I.j0 * G.gs;
        const uint32_t nl = min(I.ngroups * G.gs, G.n - first);
        for (uint32_t v = tid; v < G.n; v += kB) {
            const uint64_t r = vis[v];
            reached += (unsigned long long)__popcll(r);
            reached_arcs += (unsigned long long)__popcll(r) *
                            (off[G.node0 + v + 1] - off[G.node0 + v]);
            for (uint32_t j = 0; j < I.ngroups; j++) {
                if (j * G.gs >= nl) break;
                gsum[G.gsum_off + (size_t)(I.j0 + j) * G.n + v] =
                    group_sum(I, G, sig + (size_t)v * kLanes, r, v, j);
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

// ---- the allocation-free entry (netbc_betweenness_batch_noalloc) -----------

// Its scratch: one static device array per GPU the library runs on, held
// until the process ends, laid out as [header | node counts | graph
// descriptors | graph order | per-workgroup rows]. Host code reaches it only
// through the address hipGetSymbolAddress returns on each device.
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr uint32_t kNoAllocMaxGraphs = 1u << 16;

struct NoAllocHeader {
    uint32_t next;        // work queue: graphs handed out so far
    uint32_t bad_graph;   // least graph with an invalid row or arc (kNone: none)
    uint32_t deep_graph;  // least graph searched past 254 levels (kNone: none)
    uint32_t pad;
    unsigned long long stats[2];  // reached nodes, reached arcs
};
constexpr size_t kCountsOff = 256;
constexpr size_t kDescOff = kCountsOff + (size_t)kNoAllocMaxGraphs * sizeof(int64_t);
constexpr size_t kOrderOff = kDescOff + (size_t)kNoAllocMaxGraphs * sizeof(SmallGraph);
constexpr size_t kRowsOff = (kOrderOff + (size_t)kNoAllocMaxGraphs * 4 + 255) / 256 * 256;
static_assert(sizeof(NoAllocHeader) <= kCountsOff, "header fits its slot");
static_assert(kRowsOff < kScratchBytes, "rows region left");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Lays a batch out in the scratch at `base`: clears the header and, from
// the node counts, writes each graph's descriptor (first node, source-group
// size and count as parallel.rs:437-440, scale as betweenness.rs:100-113)
// and the order in which workgroups take the graphs (most nodes first).
// One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void noalloc_plan_kernel(unsigned char* __restrict__ base,
                                                           uint32_t graphs, int directed,
                                                           int normalized) {
    NoAllocHeader* H = reinterpret_cast<NoAllocHeader*>(base);
    const int64_t* counts = reinterpret_cast<const int64_t*>(base + kCountsOff);
    SmallGraph* desc = reinterpret_cast<SmallGraph*>(base + kDescOff);
    uint32_t* order = reinterpret_cast<uint32_t*>(base + kOrderOff);
    __shared__ uint32_t part[1024];
    __shared__ uint32_t hist[kSmallMaxNodes + 1];
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        H->next = 0;
        H->bad_graph = kNone;
        H->deep_graph = kNone;
        H->pad = 0;
        H->stats[0] = 0;
        H->stats[1] = 0;
    }
    for (uint32_t i = t; i <= kSmallMaxNodes; i += blockDim.x) hist[i] = 0;
    // graphs [g0, g1) of this thread, their first nodes by a block-wide scan
    const uint32_t per = (graphs + blockDim.x - 1) / blockDim.x;
    const uint32_t g0 = min(t * per, graphs), g1 = min(g0 + per, graphs);
    uint32_t local = 0;
    for (uint32_t g = g0; g < g1; g++) local += (uint32_t)counts[g];
    part[t] = local;
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0;
        for (uint32_t i = 0; i < blockDim.x; i++) {
            const uint32_t x = part[i];
            part[i] = run;
            run += x;
        }
    }
    __syncthreads();
    uint32_t node0 = part[t];
    for (uint32_t g = g0; g < g1; g++) {
        const uint32_t n = (uint32_t)counts[g];
        SmallGraph G;
        G.n = n;
        G.node0 = node0;
        G.gs = n ? (n + 63) / 64 : 1;
        G.groups = n ? (n + G.gs - 1) / G.gs : 0;
        G.gsum_off = 0;
        double scale = directed ? 1.0 : 0.5;
        if (normalized && n > 2) {
            const double pairs = (double)((uint64_t)(n - 1) * (uint64_t)(n - 2));
            scale *= directed ? 1.0 / pairs : 2.0 / pairs;
        }
        G.scale = scale;
        G.device = 1;
        G.pad = 0;
        desc[g] = G;
        node0 += n;
        atomicAdd(&hist[kSmallMaxNodes - n], 1u);
    }
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0;
        for (uint32_t i = 0; i <= kSmallMaxNodes; i++) {
            const uint32_t x = hist[i];
            hist[i] = run;
            run += x;
        }
    }
    __syncthreads();
    for (uint32_t g = g0; g < g1; g++)
        order[atomicAdd(&hist[kSmallMaxNodes - (uint32_t)counts[g]], 1u)] = g;
}

// Betweenness of whole graphs, one graph per workgroup at a time: its
// batches of source groups one after another (run_small_item), each group's
// sum added into the caller's scores in group order, which is the
// reference's merge bc = ((g0 + g1) + g2) + ... (parallel.rs:383-503), then
// the graph's scale. Besides the graph and the scores everything lives in
// the scratch at `base`; a workgroup's rows take row_bytes from kRowsOff on.
__global__ __launch_bounds__(kSmallBlock) void small_whole_kernel(
    unsigned char* __restrict__ base, uint32_t graphs, size_t nmax, size_t row_bytes,
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ idx,
    const uint32_t* __restrict__ in_off, const uint32_t* __restrict__ in_idx, uint64_t arcs,
    double* __restrict__ out) {
    __shared__ uint64_t vis[kSmallMaxNodes];
    __shared__ uint64_t fa[kSmallMaxNodes];
    __shared__ uint64_t fb[kSmallMaxNodes];
    __shared__ uint32_t item_s;
    NoAllocHeader* H = reinterpret_cast<NoAllocHeader*>(base);
    const SmallGraph* desc = reinterpret_cast<const SmallGraph*>(base + kDescOff);
    const uint32_t* order = reinterpret_cast<const uint32_t*>(base + kOrderOff);
    unsigned char* rows = base + kRowsOff + (size_t)blockIdx.x * row_bytes;
    double* sig = reinterpret_cast<double*>(rows);
    double* coef = sig + nmax * kLanes;
    uint8_t* lvl = reinterpret_cast<uint8_t*>(coef + nmax * kLanes);
    const int tid = threadIdx.x, lane = tid & 63;
    const ArcCheck chk{arcs, &H->bad_graph};
    unsigned long long reached = 0, reached_arcs = 0;
    for (;;) {
        if (tid == 0) item_s = atomicAdd(&H->next, 1u);
        __syncthreads();
        const uint32_t it = item_s;
        __syncthreads();
        if (it >= graphs) break;
        const uint32_t g = order[it];
        const SmallGraph G = desc[g];
        const uint32_t per = max(1u, 64u / G.gs);
        bool ok = true;
        for (uint32_t j0 = 0; j0 < G.groups; j0 += per) {
            const SmallItem I{g, j0, min(per, G.groups - j0), 0};
            ok = run_small_item<double, true>(I, G, off, idx, in_off, in_idx, sig, coef, lvl, vis,
                                              fa, fb, chk);
            if (!ok) {
                if (tid == 0) atomicMin(&H->deep_graph, g);
                break;
            }
            const uint32_t first = j0 * G.gs;
            const uint32_t nl = min(I.ngroups * G.gs, G.n - first);
            for (uint32_t v = tid; v < G.n; v += kSmallBlock) {
                const uint64_t r = vis[v];
                reached += (unsigned long long)__popcll(r);
                reached_arcs += (unsigned long long)__popcll(r) *
                                (off[G.node0 + v + 1] - off[G.node0 + v]);
                double* o = &out[G.node0 + v];
                for (uint32_t j = 0; j < I.ngroups; j++) {
                    if (j * G.gs >= nl) break;
                    const double acc = group_sum(I, G, sig + (size_t)v * kLanes, r, v, j);
                    *o = j0 + j == 0 ? acc : *o + acc;
                }
            }
            __syncthreads();
        }
        // each score was written by this thread in the loop above
        if (ok && G.scale != 1.0)
            for (uint32_t v = tid; v < G.n; v += kSmallBlock) out[G.node0 + v] *= G.scale;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        reached += (unsigned long long)__shfl_xor((long long)reached, o, 64);
        reached_arcs += (unsigned long long)__shfl_xor((long long)reached_arcs, o, 64);
    }
    if (lane == 0 && reached) atomicAdd(&H->stats[0], reached);
    if (lane == 0 && reached_arcs) atomicAdd(&H->stats[1], reached_arcs);
}

// The scratch's address on each device (looked up once per device, with
// that device current) and how many whole-graph workgroups fit on it.
struct ScratchDevice {
    bool ready = false;
    unsigned char* ptr = nullptr;
    unsigned resident = 1;
};
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
ScratchDevice g_scratch_dev[kMaxDevices];

ScratchDevice scratch_on_current_device() {
    int dev = 0;
    NETBC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error("device index beyond the library's table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    ScratchDevice& s = g_scratch_dev[dev];
    if (!s.ready) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(std::string("device error: ") + hipGetErrorString(e) +
                        " (the address of the allocation-free scratch)");
        int cus = 0, per = 0;
        NETBC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        NETBC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, small_whole_kernel,
                                                               kSmallBlock, 0));
        s.ptr = static_cast<unsigned char*>(p);
        s.resident = (unsigned)std::max(1, cus * std::max(per, 1));
        s.ready = true;
    }
    return s;
}

// The coefficient type of a small_batch_kernel instance (for dispatch).
template <class C>
C* kernel_coef(void (*)(const SmallItem*, uint32_t, uint32_t*, const SmallGraph*,
                        const uint32_t*, const uint32_t*, const uint32_t*, const uint32_t*,
                        double*, C*, uint8_t*, size_t, double*, uint32_t*,
                        unsigned long long*)) {
    return nullptr;
}

// Calls f with the small_batch_kernel instance for the precision and
// workgroup size (256, 512 or 1024 threads).
template <class F>
void with_small_kernel(int precision, int block, F&& f) {
    if (precision) {
        if (block == 256) f(small_batch_kernel<float, 256>);
        else if (block == 1024) f(small_batch_kernel<float, 1024>);
        else f(small_batch_kernel<float, 512>);
    } else {
        if (block == 256) f(small_batch_kernel<double, 256>);
        else if (block == 1024) f(small_batch_kernel<double, 1024>);
        else f(small_batch_kernel<double, 512>);
    }
}

}  // namespace

void small_setup(SmallRun& r, const std::vector<SmallGraph>& graphs,
                 const std::vector<uint32_t>& node_graph, const RunOptions& opt, hipStream_t st) {
    r.total = node_graph.size();
    r.ngraphs = graphs.size();
    r.precision = opt.precision;
    // Work items, largest graphs first.
    std::vector<uint32_t> order(graphs.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = (uint32_t)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return graphs[a].n > graphs[b].n; });
    std::vector<SmallItem> items;
    size_t gsum_total = 0;
    uint32_t nmax = 1;
    for (const SmallGraph& G : graphs)
        gsum_total = std::max<size_t>(gsum_total, G.gsum_off + (size_t)G.groups * G.n);
    for (uint32_t gi : order) {
        const SmallGraph& G = graphs[gi];
        if (G.n == 0 || !G.device) continue;
        nmax = std::max(nmax, G.n);
        uint32_t per = std::max<uint32_t>(1, 64 / G.gs);
        for (uint32_t j = 0; j < G.groups; j += per)
            items.push_back(SmallItem{gi, j, std::min(per, G.groups - j), 0});
    }
    r.nitems = (uint32_t)items.size();
    r.nmax = nmax;
    r.graphs.alloc(graphs.size());
    r.node_graph.alloc(r.total);
    r.too_deep.alloc(graphs.size());
    r.counter.alloc(1);
    r.items.alloc(items.size());
    r.gsum.alloc(gsum_total);
    r.out.alloc(r.total);
    r.stats.alloc(2);
    NETBC_HIP(hipMemcpyAsync(r.graphs.p, graphs.data(), graphs.size() * sizeof(SmallGraph),
                             hipMemcpyHostToDevice, st));
    NETBC_HIP(hipMemcpyAsync(r.node_graph.p, node_graph.data(), r.total * 4,
                             hipMemcpyHostToDevice, st));
    if (!items.empty())
        NETBC_HIP(hipMemcpyAsync(r.items.p, items.data(), items.size() * sizeof(SmallItem),
                                 hipMemcpyHostToDevice, st));

    int dev = 0;
    NETBC_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    NETBC_HIP(hipGetDeviceProperties(&prop, dev));
    // workgroup size: the caller's, else the tuning table's class 2 (many
    // small unweighted graphs), else kSmallBlock
    r.block = kSmallBlock;
    for (const TuneEntry& e : kTuning)
        if (e.kind == 2) {
            r.block = e.sweep_block;
            break;
        }
    if (opt.sweep_block > 0) r.block = opt.sweep_block;
    if (r.block != 256 && r.block != 512 && r.block != 1024)
        throw Error("the one-workgroup engine runs workgroups of 256, 512 or 1024 threads, not " +
                    std::to_string(r.block));
    int per_cu = 0;
    with_small_kernel(r.precision, r.block, [&](auto kernel) {
        NETBC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, r.block, 0));
    });
    per_cu = std::max(1, per_cu);
    size_t grid = std::min<size_t>((size_t)prop.multiProcessorCount * per_cu,
                                   std::max<size_t>(items.size(), 1));
    if (opt.slots > 0) grid = std::min<size_t>(grid, (size_t)opt.slots);
    r.grid = grid;
    const size_t per = (size_t)nmax * kLanes;
    r.sig.alloc(grid * per);
    r.lvl.alloc(grid * per);
    if (opt.precision) r.coef32.alloc(grid * per);
    else r.coef64.alloc(grid * per);
    // the uploads read host vectors that end with this call
    NETBC_HIP(hipStreamSynchronize(st));
}

void small_enqueue(SmallRun& r, const uint32_t* off, const uint32_t* idx, const uint32_t* in_off,
                   const uint32_t* in_idx, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    // Clearing by kernels rather than memsets, so that a captured graph
    // clears on every replay. gsum needs none: every group sum of a graph
    // the kernel takes is written, and the scores of the other graphs
    // (too deep, or not taken) are recomputed by the caller.
    zero_u32_kernel<<<1, 64, 0, st>>>(r.counter.p, 1);
    NETBC_HIP(hipGetLastError());
    zero_u32_kernel<<<(unsigned)((r.ngraphs + 255) / 256 + 1), 256, 0, st>>>(r.too_deep.p,
                                                                           r.ngraphs);
    NETBC_HIP(hipGetLastError());
    zero_u32_kernel<<<1, 64, 0, st>>>(reinterpret_cast<uint32_t*>(r.stats.p), 4);
    NETBC_HIP(hipGetLastError());
    if (e0) NETBC_HIP(hipEventRecord(e0, st));
    if (r.nitems) {
        with_small_kernel(r.precision, r.block, [&](auto kernel) {
            using C = std::remove_pointer_t<decltype(kernel_coef(kernel))>;
            C* coef = nullptr;
            if constexpr (std::is_same_v<C, float>) coef = r.coef32.p;
            else coef = r.coef64.p;
            kernel<<<(unsigned)r.grid, r.block, 0, st>>>(
                r.items.p, r.nitems, r.counter.p, r.graphs.p, off, idx, in_off, in_idx, r.sig.p,
                coef, r.lvl.p, r.nmax, r.gsum.p, r.too_deep.p, r.stats.p);
        });
        NETBC_HIP(hipGetLastError());
    }
    if (e1) NETBC_HIP(hipEventRecord(e1, st));
    if (r.total) {
        small_reduce_kernel<<<(unsigned)((r.total + 255) / 256), 256, 0, st>>>(
            r.graphs.p, r.node_graph.p, r.gsum.p, r.total, r.out.p);
        NETBC_HIP(hipGetLastError());
    }
}

void small_fetch(SmallRun& r, double* out_host, std::vector<uint32_t>& too_deep,
                 unsigned long long stats[2], hipStream_t st) {
    if (r.total)
        NETBC_HIP(hipMemcpyAsync(out_host, r.out.p, r.total * 8, hipMemcpyDeviceToHost, st));
    too_deep.assign(r.ngraphs, 0);
    if (r.ngraphs)
        NETBC_HIP(hipMemcpyAsync(too_deep.data(), r.too_deep.p, r.ngraphs * 4,
                                 hipMemcpyDeviceToHost, st));
    NETBC_HIP(hipMemcpyAsync(stats, r.stats.p, 16, hipMemcpyDeviceToHost, st));
    NETBC_HIP(hipStreamSynchronize(st));
}

void run_small_batch(Graph& g, const std::vector<SmallGraph>& graphs,
                     const std::vector<uint32_t>& node_graph, const RunOptions& opt,
                     double* out_host, std::vector<uint32_t>& too_deep, RunDiag& diag) {
    hipStream_t st = g.stream;
    const DeviceGraph& d = g.dev;
    SmallRun r;
    small_setup(r, graphs, node_graph, opt, st);
    hipEvent_t e0, e1;
    NETBC_HIP(hipEventCreate(&e0));
    NETBC_HIP(hipEventCreate(&e1));
    unsigned long long hs[2] = {0, 0};
    try {
        small_enqueue(r, d.out_off.p, d.out_idx.p, d.in_off_p(), d.in_idx_p(), st, e0, e1);
        small_fetch(r, out_host, too_deep, hs, st);
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    float ms = 0.0f;
    NETBC_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    diag.kernel_ms += ms;
    diag.dominant_ms += ms;
    diag.dominant_launches += 1;
    diag.reached_nodes += (int64_t)hs[0];
    diag.reached_arcs += (int64_t)hs[1];
    diag.slots = (int32_t)r.grid;
}

}  // namespace netbc

namespace netbc {

void run_small_batch_noalloc(int64_t graphs, const int64_t* node_counts, bool directed,
                             bool normalized, const uint32_t* off, const uint32_t* idx,
                             const uint32_t* in_off, const uint32_t* in_idx, int64_t arcs,
                             double* out, hipStream_t st, RunDiag& diag) {
    if (graphs < 0) throw Error("negative graph count");
    if ((uint64_t)graphs > kNoAllocMaxGraphs)
        throw Error("the allocation-free entry takes at most " +
                    std::to_string(kNoAllocMaxGraphs) + " graphs per call");
    if (graphs > 0 && !node_counts) throw Error("null node counts");
    uint64_t total = 0, nmax = 0;
    for (int64_t i = 0; i < graphs; i++) {
        if (node_counts[i] < 0) throw Error("graph " + std::to_string(i) + ": negative node count");
        if (node_counts[i] > (int64_t)kSmallMaxNodes)
            throw Error("graph " + std::to_string(i) + ": " + std::to_string(node_counts[i]) +
                        " nodes; the allocation-free entry takes graphs of at most " +
                        std::to_string(kSmallMaxNodes) + " nodes");
        total += (uint64_t)node_counts[i];
        nmax = std::max<uint64_t>(nmax, (uint64_t)node_counts[i]);
    }
    if (total == 0) return;
    if (!off || !idx || !out) throw Error("null offsets, targets or scores");
    if (directed && (!in_off || !in_idx))
        throw Error("a directed batch needs its in-CSR (in_offsets, in_targets)");
    if (!directed) {
        in_off = off;
        in_idx = idx;
    }
    if (arcs < 0 || (uint64_t)arcs > 0xFFFFFFFEull) throw Error("arc count outside 0..2^32 - 2");
    check_device_range(out, total * sizeof(double), "score buffer");
    check_device_range(off, (total + 1) * sizeof(uint32_t), "offset array");
    if (arcs) check_device_range(idx, (size_t)arcs * sizeof(uint32_t), "target array");
    if (directed) {
        check_device_range(in_off, (total + 1) * sizeof(uint32_t), "in-offset array");
        if (arcs) check_device_range(in_idx, (size_t)arcs * sizeof(uint32_t), "in-target array");
    }
    const size_t row_bytes = ((size_t)nmax * kLanes * (8 + 8 + 1) + 255) / 256 * 256;
    const size_t fit = (kScratchBytes - kRowsOff) / row_bytes;
    if (fit == 0) throw Error("the batch's rows exceed the allocation-free scratch");
    const ScratchDevice sd = scratch_on_current_device();
    const unsigned grid =
        (unsigned)std::min<uint64_t>({(uint64_t)fit, (uint64_t)sd.resident, (uint64_t)graphs});
    NETBC_HIP(hipMemcpyAsync(sd.ptr + kCountsOff, node_counts, (size_t)graphs * sizeof(int64_t),
                             hipMemcpyHostToDevice, st));
    noalloc_plan_kernel<<<1, 1024, 0, st>>>(sd.ptr, (uint32_t)graphs, directed ? 1 : 0,
                                            normalized ? 1 : 0);
    NETBC_HIP(hipGetLastError());
    small_whole_kernel<<<grid, kSmallBlock, 0, st>>>(sd.ptr, (uint32_t)graphs, nmax, row_bytes,
                                                     off, idx, in_off, in_idx, (uint64_t)arcs,
                                                     out);
    NETBC_HIP(hipGetLastError());
    NoAllocHeader h{};
    NETBC_HIP(hipMemcpyAsync(&h, sd.ptr, sizeof h, hipMemcpyDeviceToHost, st));
    NETBC_HIP(hipStreamSynchronize(st));
    if (h.bad_graph != kNone)
        throw Error("graph " + std::to_string(h.bad_graph) +
                    ": an offset runs backwards or past the arc count, or an arc leaves the "
                    "graph's nodes");
    if (h.deep_graph != kNone)
        throw Error("graph " + std::to_string(h.deep_graph) +
                    ": a search passed 254 levels; the allocation-free entry does not take "
                    "such graphs (netbc_betweenness_batch does)");
    diag.reached_nodes += (int64_t)h.stats[0];
    diag.reached_arcs += (int64_t)h.stats[1];
    diag.slots = (int32_t)grid;
}

}  // namespace netbc
