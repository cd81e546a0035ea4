// Batched Brandes betweenness on one GPU.
//
// Sources are processed 64 at a time: one batch = one 64-bit lane mask per
// node, and lane l of a wavefront owns source l of the batch. Per-lane
// values are stored as rows of 64 (node-major), so a wavefront touching a
// node reads or writes its whole row in one coalesced access. Several
// batches ("slots") run together in the same launches.
//
// Per source the arithmetic is the reference's
// (crates/engine/src/metrics/betweenness.rs:161-295):
//   forward   sigma[w] = sum of sigma[v] over arcs v->w on shortest paths
//             (exact integers in FP64 below 2^53, so the pull order used
//             here gives the same bits as the reference's push order);
//   backward  dv = sum of coef[w] over successors w, in CSR row order,
//             delta = sigma[v] * dv, coef[v] = (1 + delta) / sigma[v];
//   sum       bc[v] += delta for v != s, sources summed in 64 contiguous
//             groups merged in group order (parallel.rs:383-503).
// The file is compiled with -ffp-contract=off so no multiply-add is fused
// and every FP64 result matches the reference's bit for bit.

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>

#include "netbc_internal.h"

namespace netbc {
namespace {

struct QEnt {
    uint32_t v;
    uint32_t b;
    unsigned long long m;
};

// Per-super-batch counters come first (cleared before every super-batch);
// the rest accumulate over the whole run.
struct Ctrl {
    uint32_t qlen;
    uint32_t tlen[2];
    uint32_t qoverflow;
    uint32_t pending;
    uint32_t pad;
    uint32_t sig_overflow;
    uint32_t pad2;
    unsigned long long max_sigma_bits;
    unsigned long long reached_nodes;
    unsigned long long reached_arcs;
    unsigned long long fwd_arcs;  // arcs scanned by top-down expansion
};

constexpr double kTwo53 = 9007199254740992.0;

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ size_t global_wave() {
    return ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
}
__device__ __forceinline__ size_t wave_count() { return ((size_t)gridDim.x * blockDim.x) >> 6; }
__device__ __forceinline__ uint32_t rl32(uint32_t x, int j) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, j);
}
__device__ __forceinline__ uint64_t rl64(uint64_t x, int j) {
    uint32_t lo = rl32((uint32_t)x, j), hi = rl32((uint32_t)(x >> 32), j);
    return (uint64_t)hi << 32 | lo;
}
__device__ __forceinline__ uint64_t lanemask_lt(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Largest value across the wavefront (doubles are >= 0 here).
__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
    return x;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__device__ __forceinline__ void note_sigma(Ctrl* ctrl, double s, bool active) {
    double mx = wave_max(active ? s : 0.0);
    if (lane_id() == 0 && mx > 0.0) {
        unsigned long long bits = (unsigned long long)__double_as_longlong(mx);
        atomicMax(&ctrl->max_sigma_bits, bits);
        if (mx > kTwo53) ctrl->sig_overflow = 1;
    }
}

// ---- initialisation -------------------------------------------------------

// Marks the source of every active lane: visited, sigma = 1, level 0.
__global__ void init_sources_kernel(const int64_t* __restrict__ list, int64_t item0, int64_t k,
                                    int slots, size_t n, uint64_t* __restrict__ vis,
                                    uint64_t* __restrict__ g0, double* __restrict__ sig,
                                    QEnt* __restrict__ q, Ctrl* ctrl, uint32_t* __restrict__ dist) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= slots * kLanes) return;
    int b = t / kLanes, l = t % kLanes;
    int64_t i = item0 + t;
    if (i >= k) return;
    uint32_t s = (uint32_t)(list ? list[i] : i);
    unsigned long long bit = 1ull << l;
    size_t mi = (size_t)b * n + s;
    atomicOr((unsigned long long*)&vis[mi], bit);
    if (g0) atomicOr((unsigned long long*)&g0[mi], bit);
    sig[mi * kLanes + l] = 1.0;
    if (dist) dist[mi * kLanes + l] = 0;
    uint32_t pos = atomicAdd(&ctrl->qlen, 1u);
    q[pos] = QEnt{s, (uint32_t)b, bit};
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, size_t count, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = v;
}

__global__ void fill_u64_kernel(uint64_t* __restrict__ p, size_t count, uint64_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = v;
}

// ---- breadth-first search: forward -----------------------------------------

// Top-down step of level d: every frontier entry (v, b, M) offers its lanes
// to each out-neighbour x that has not yet been reached in them. The first
// offer to x in this level records x in the touched list. Threads also
// clear the stale level d-1 entries of the next-level bitmap.
__global__ void bfs_expand_kernel(const QEnt* __restrict__ q, uint32_t qbeg, uint32_t qend,
                                  uint32_t zbeg, uint32_t zend, const uint32_t* __restrict__ off,
                                  const uint32_t* __restrict__ idx,
                                  const uint64_t* __restrict__ vis, uint64_t* __restrict__ cand,
                                  uint64_t* __restrict__ nxt, uint2* __restrict__ touched,
                                  uint32_t* __restrict__ tlen, size_t n) {
    const int lane = lane_id();
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nth = (size_t)gridDim.x * blockDim.x;
    for (size_t i = zbeg + tid; i < zend; i += nth) {
        QEnt e = q[i];
        nxt[(size_t)e.b * n + e.v] = 0;
    }
    const uint64_t lt = lanemask_lt(lane);
    for (size_t e = qbeg + global_wave(); e < qend; e += wave_count()) {
        QEnt en = q[e];
        size_t base = (size_t)en.b * n;
        uint32_t a0 = off[en.v], a1 = off[en.v + 1];
        for (uint32_t c = a0; c < a1; c += 64) {
            uint32_t a = c + lane;
            bool ok = a < a1;
            uint32_t x = ok ? idx[a] : 0;
            uint64_t nv = ok ? (en.m & ~vis[base + x]) : 0;
            bool first = false;
            if (nv) {
                unsigned long long old =
                    atomicOr((unsigned long long*)&cand[base + x], (unsigned long long)nv);
                first = old == 0;
            }
            uint64_t bal = __ballot(first);
            if (bal) {
                int leader = __ffsll((unsigned long long)bal) - 1;
                uint32_t pos = 0;
                if (lane == leader) pos = atomicAdd(tlen, (uint32_t)__popcll(bal));
                pos = rl32(pos, leader);
                if (first) touched[pos + __popcll(bal & lt)] = make_uint2(x, en.b);
            }
        }
    }
}

// Path counts of the nodes first reached at level d+1: for each touched
// node x with new lanes N, sigma[x][l] = sum over in-arcs u->x with u at
// level d in lane l of sigma[u][l]. Appends (x, b, N) to the level queue.
__global__ void bfs_sigma_kernel(const uint2* __restrict__ touched,
                                 const uint32_t* __restrict__ tlen_cur,
                                 uint32_t* __restrict__ tlen_next,
                                 const uint32_t* __restrict__ in_off,
                                 const uint32_t* __restrict__ in_idx,
                                 const uint64_t* __restrict__ cur, uint64_t* __restrict__ cand,
                                 uint64_t* __restrict__ vis, uint64_t* __restrict__ nxt,
                                 double* __restrict__ sig, QEnt* __restrict__ q, uint32_t qcap,
                                 Ctrl* ctrl, size_t n) {
    const int lane = lane_id();
    if (blockIdx.x == 0 && threadIdx.x == 0) *tlen_next = 0;
    const uint32_t T = *tlen_cur;
    for (size_t t = global_wave(); t < T; t += wave_count()) {
        uint2 te = touched[t];
        uint32_t x = te.x;
        size_t base = (size_t)te.y * n;
        uint64_t N = cand[base + x];
        double acc = 0.0;
        uint32_t a0 = in_off[x], a1 = in_off[x + 1];
        for (uint32_t c = a0; c < a1; c += 64) {
            uint32_t a = c + lane;
            bool ok = a < a1;
            uint32_t u = ok ? in_idx[a] : 0;
            uint64_t mm = ok ? (cur[base + u] & N) : 0;
            uint64_t bal = __ballot(mm != 0);
            while (bal) {
                int j = __ffsll((unsigned long long)bal) - 1;
                bal &= bal - 1;
                uint32_t uu = rl32(u, j);
                uint64_t m2 = rl64(mm, j);
                if ((m2 >> lane) & 1) acc += sig[(base + uu) * kLanes + lane];
            }
        }
        bool mine = (N >> lane) & 1;
        if (mine) sig[(base + x) * kLanes + lane] = acc;
        note_sigma(ctrl, acc, mine);
        if (lane == 0) {
            vis[base + x] |= N;
            nxt[base + x] = N;
            cand[base + x] = 0;
            uint32_t pos = atomicAdd(&ctrl->qlen, 1u);
            if (pos < qcap)
                q[pos] = QEnt{x, te.y, (unsigned long long)N};
            else
                ctrl->qoverflow = 1;
        }
    }
}

// ---- breadth-first search: backward ---------------------------------------

template <class C>
struct Dep;

template <>
struct Dep<double> {
    double s = 0.0;
    __device__ __forceinline__ void add(double x) { s += x; }
    __device__ __forceinline__ double value() const { return s; }
    __device__ __forceinline__ static double coef(double delta, double sigma) {
        return (1.0 + delta) / sigma;
    }
};

// FP32 dependencies with Kahan-compensated successor sums.
template <>
struct Dep<float> {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float x) {
        float y = x - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
    __device__ __forceinline__ static float coef(double delta, double sigma) {
        return (float)((1.0 + delta) / sigma);
    }
};

// Dependencies of the level-d entries. R is the level d+1 bitmap, W the
// bitmap being turned from level d+2 into level d (lanes are disjoint, so
// the OR and AND-NOT updates commute).
template <class C>
__global__ void bfs_backward_kernel(const QEnt* __restrict__ q, uint32_t qbeg, uint32_t qend,
                                    uint32_t abeg, uint32_t aend, int last_level,
                                    const uint32_t* __restrict__ off,
                                    const uint32_t* __restrict__ idx,
                                    const uint64_t* __restrict__ R, uint64_t* __restrict__ W,
                                    double* __restrict__ sig, C* __restrict__ coef, size_t n) {
    const int lane = lane_id();
    if (!last_level) {
        size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
        size_t nth = (size_t)gridDim.x * blockDim.x;
        for (size_t i = qbeg + tid; i < qend; i += nth) {
            QEnt e = q[i];
            atomicOr((unsigned long long*)&W[(size_t)e.b * n + e.v], e.m);
        }
        for (size_t i = abeg + tid; i < aend; i += nth) {
            QEnt e = q[i];
            atomicAnd((unsigned long long*)&W[(size_t)e.b * n + e.v], ~e.m);
        }
    }
    for (size_t e = qbeg + global_wave(); e < qend; e += wave_count()) {
        QEnt en = q[e];
        size_t base = (size_t)en.b * n;
        uint64_t M = en.m;
        Dep<C> dv;
        if (!last_level) {
            uint32_t a0 = off[en.v], a1 = off[en.v + 1];
            for (uint32_t c = a0; c < a1; c += 64) {
                uint32_t a = c + lane;
                bool ok = a < a1;
                uint32_t x = ok ? idx[a] : 0;
                uint64_t mm = ok ? (R[base + x] & M) : 0;
                uint64_t bal = __ballot(mm != 0);
                while (bal) {
                    int j = __ffsll((unsigned long long)bal) - 1;
                    bal &= bal - 1;
                    uint32_t xx = rl32(x, j);
                    uint64_t m2 = rl64(mm, j);
                    if ((m2 >> lane) & 1) dv.add(coef[(base + xx) * kLanes + lane]);
                }
            }
        }
        if ((M >> lane) & 1) {
            size_t r = (base + en.v) * kLanes + lane;
            double s = sig[r];
            double delta = s * dv.value();
            coef[r] = Dep<C>::coef(delta, s);
            sig[r] = delta;
        }
    }
}

// ---- integer lengths (Dial / delta-stepping with unit buckets) -------------

// Drains bucket k at distance D: the lanes of x whose tentative distance is
// D are final. Their path counts are pulled over in-arcs u->x of length
// len with dist[u] + len == D, and (x, b, M) joins the level queue.
__global__ void dial_settle_kernel(const uint2* __restrict__ list, const uint32_t* __restrict__ blen,
                                   size_t list_cap, uint32_t D, int k, int words,
                                   uint64_t* __restrict__ inb,
                                   const uint32_t* __restrict__ in_off,
                                   const uint32_t* __restrict__ in_idx,
                                   const uint32_t* __restrict__ in_len,
                                   const uint32_t* __restrict__ dist, uint64_t* __restrict__ vis,
                                   double* __restrict__ sig, QEnt* __restrict__ q, uint32_t qcap,
                                   Ctrl* ctrl, size_t n) {
    const int lane = lane_id();
    const uint32_t T = min(blen[k], (uint32_t)list_cap);
    for (size_t t = global_wave(); t < T; t += wave_count()) {
        uint2 te = list[t];
        uint32_t x = te.x;
        size_t base = (size_t)te.y * n;
        if (lane == 0)
            atomicAnd((unsigned long long*)&inb[(base + x) * words + (k >> 6)],
                      ~(1ull << (k & 63)));
        uint64_t M = __ballot(dist[(base + x) * kLanes + lane] == D);
        if (M == 0) continue;
        double acc = 0.0;
        uint32_t a0 = in_off[x], a1 = in_off[x + 1];
        for (uint32_t c = a0; c < a1; c += 64) {
            uint32_t a = c + lane;
            bool ok = a < a1;
            uint32_t u = ok ? in_idx[a] : 0;
            uint32_t len = ok ? in_len[a] : 0;
            uint64_t mm = ok ? (vis[base + u] & M) : 0;
            uint64_t bal = __ballot(mm != 0);
            while (bal) {
                int j = __ffsll((unsigned long long)bal) - 1;
                bal &= bal - 1;
                uint32_t uu = rl32(u, j);
                uint32_t ll = rl32(len, j);
                uint64_t m2 = rl64(mm, j);
                if ((m2 >> lane) & 1) {
                    size_t r = (base + uu) * kLanes + lane;
                    uint32_t du = dist[r];
                    if ((uint64_t)du + ll == D) acc += sig[r];
                }
            }
        }
        bool mine = (M >> lane) & 1;
        if (mine) sig[(base + x) * kLanes + lane] = acc;
        note_sigma(ctrl, acc, mine);
        if (lane == 0) {
            vis[base + x] |= M;
            uint32_t pos = atomicAdd(&ctrl->qlen, 1u);
            if (pos < qcap)
                q[pos] = QEnt{x, te.y, (unsigned long long)M};
            else
                ctrl->qoverflow = 1;
        }
    }
}

// Relaxes the out-arcs of the entries settled at distance D. A lane whose
// tentative distance drops puts its node into bucket (D + len) mod W once.
__global__ void dial_relax_kernel(const QEnt* __restrict__ q, uint32_t qbeg, uint32_t qend,
                                  uint32_t D, int W, int words, int k_drained,
                                  uint32_t* __restrict__ blen, uint2* __restrict__ lists,
                                  size_t list_cap, uint64_t* __restrict__ inb,
                                  const uint32_t* __restrict__ off,
                                  const uint32_t* __restrict__ idx,
                                  const uint32_t* __restrict__ len_arr,
                                  uint32_t* __restrict__ dist, const uint64_t* __restrict__ vis,
                                  Ctrl* ctrl, size_t n) {
    const int lane = lane_id();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicSub(&ctrl->pending, blen[k_drained]);
        blen[k_drained] = 0;
    }
    for (size_t e = qbeg + global_wave(); e < qend; e += wave_count()) {
        QEnt en = q[e];
        size_t base = (size_t)en.b * n;
        uint64_t M = en.m;
        uint32_t a0 = off[en.v], a1 = off[en.v + 1];
        for (uint32_t c = a0; c < a1; c += 64) {
            uint32_t a = c + lane;
            bool ok = a < a1;
            uint32_t y = ok ? idx[a] : 0;
            uint32_t len = ok ? len_arr[a] : 0;
            uint64_t mm = ok ? (M & ~vis[base + y]) : 0;
            uint64_t bal = __ballot(mm != 0);
            while (bal) {
                int j = __ffsll((unsigned long long)bal) - 1;
                bal &= bal - 1;
                uint32_t yy = rl32(y, j);
                uint32_t nd = D + rl32(len, j);
                uint64_t m2 = rl64(mm, j);
                bool improved = false;
                if ((m2 >> lane) & 1) {
                    uint32_t* p = &dist[(base + yy) * kLanes + lane];
                    if (nd < *p) improved = atomicMin(p, nd) > nd;
                }
                if (__ballot(improved) && lane == 0) {
                    int kk = (int)(nd % (uint32_t)W);
                    unsigned long long bit = 1ull << (kk & 63);
                    unsigned long long old = atomicOr(
                        (unsigned long long*)&inb[(base + yy) * words + (kk >> 6)], bit);
                    if (!(old & bit)) {
                        uint32_t pos = atomicAdd(&blen[kk], 1u);
                        atomicAdd(&ctrl->pending, 1u);
                        if (pos < list_cap)
                            lists[(size_t)kk * list_cap + pos] = make_uint2(yy, en.b);
                        else
                            ctrl->qoverflow = 2;
                    }
                }
            }
        }
    }
}

template <class C>
__global__ void dial_backward_kernel(const QEnt* __restrict__ q, uint32_t qbeg, uint32_t qend,
                                     uint32_t D, const uint32_t* __restrict__ off,
                                     const uint32_t* __restrict__ idx,
                                     const uint32_t* __restrict__ len_arr,
                                     const uint32_t* __restrict__ dist,
                                     const uint64_t* __restrict__ vis, double* __restrict__ sig,
                                     C* __restrict__ coef, size_t n) {
    const int lane = lane_id();
    for (size_t e = qbeg + global_wave(); e < qend; e += wave_count()) {
        QEnt en = q[e];
        size_t base = (size_t)en.b * n;
        uint64_t M = en.m;
        Dep<C> dv;
        uint32_t a0 = off[en.v], a1 = off[en.v + 1];
        for (uint32_t c = a0; c < a1; c += 64) {
            uint32_t a = c + lane;
            bool ok = a < a1;
            uint32_t x = ok ? idx[a] : 0;
            uint32_t len = ok ? len_arr[a] : 0;
            uint64_t mm = ok ? (vis[base + x] & M) : 0;
            uint64_t bal = __ballot(mm != 0);
            while (bal) {
                int j = __ffsll((unsigned long long)bal) - 1;
                bal &= bal - 1;
                uint32_t xx = rl32(x, j);
                uint64_t want = (uint64_t)D + rl32(len, j);
                uint64_t m2 = rl64(mm, j);
                if ((m2 >> lane) & 1) {
                    size_t r = (base + xx) * kLanes + lane;
                    if ((uint64_t)dist[r] == want) dv.add(coef[r]);
                }
            }
        }
        if ((M >> lane) & 1) {
            size_t r = (base + en.v) * kLanes + lane;
            double s = sig[r];
            double delta = s * dv.value();
            coef[r] = Dep<C>::coef(delta, s);
            sig[r] = delta;
        }
    }
}

// ---- real lengths: the reference's Dijkstra, one source per thread ---------

struct DjkState {
    double* dist;
    double* sigma;
    double* coef;
    uint32_t* rank;
    uint32_t* order;
    double* hkey;
    uint32_t* hnode;
    uint32_t* pos;
};

__device__ __forceinline__ bool h_less(double da, uint32_t va, double db, uint32_t vb) {
    return da < db || (da == db && va < vb);
}

// Dijkstra with the reference's 4-ary indexed heap ordered by (distance,
// node) (weights.rs:155-277) and ties within a relative 1e-10
// (betweenness.rs:117-120, 297-368). Each thread owns private arrays.
__global__ void dijkstra_kernel(const int64_t* __restrict__ list, int64_t item0, int64_t k,
                                int slots, size_t n, const uint32_t* __restrict__ off,
                                const uint32_t* __restrict__ idx, const double* __restrict__ lens,
                                DjkState st, uint64_t* __restrict__ vis, double* __restrict__ sig,
                                Ctrl* ctrl) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= slots * kLanes) return;
    int64_t item = item0 + t;
    if (item >= k) return;
    int b = t / kLanes, l = t % kLanes;
    uint32_t s = (uint32_t)(list ? list[item] : item);
    size_t o = (size_t)t * n;
    double* dist = st.dist + o;
    double* sigma = st.sigma + o;
    double* coef = st.coef + o;
    uint32_t* rank = st.rank + o;
    uint32_t* order = st.order + o;
    double* hk = st.hkey + o;
    uint32_t* hv = st.hnode + o;
    uint32_t* pos = st.pos + o;
    for (size_t v = 0; v < n; v++) {
        dist[v] = INFINITY;
        sigma[v] = 0.0;
        coef[v] = 0.0;
        rank[v] = kNone;
        pos[v] = kNone;
    }
    size_t hlen = 0, norder = 0;
    auto sift_up = [&](size_t i) {
        double kd = hk[i];
        uint32_t kv = hv[i];
        while (i > 0) {
            size_t p = (i - 1) / 4;
            if (!h_less(kd, kv, hk[p], hv[p])) break;
            hk[i] = hk[p];
            hv[i] = hv[p];
            pos[hv[i]] = (uint32_t)i;
            i = p;
        }
        hk[i] = kd;
        hv[i] = kv;
        pos[kv] = (uint32_t)i;
    };
    auto sift_down = [&](size_t i) {
        double kd = hk[i];
        uint32_t kv = hv[i];
        for (;;) {
            size_t first = i * 4 + 1;
            if (first >= hlen) break;
            size_t best = first;
            size_t end = first + 4 < hlen ? first + 4 : hlen;
            for (size_t c = first + 1; c < end; c++)
                if (h_less(hk[c], hv[c], hk[best], hv[best])) best = c;
            if (!h_less(hk[best], hv[best], kd, kv)) break;
            hk[i] = hk[best];
            hv[i] = hv[best];
            pos[hv[i]] = (uint32_t)i;
            i = best;
        }
        hk[i] = kd;
        hv[i] = kv;
        pos[kv] = (uint32_t)i;
    };
    auto push_or_decrease = [&](uint32_t node, double d) {
        uint32_t p = pos[node];
        size_t i;
        if (p == kNone) {
            hk[hlen] = d;
            hv[hlen] = node;
            i = hlen++;
        } else {
            i = p;
            if (!h_less(d, node, hk[i], hv[i])) return;
            hk[i] = d;
        }
        sift_up(i);
    };
    auto approx_eq = [](double a, double c) {
        return fabs(a - c) <= 1e-10 * fmax(fabs(a), fabs(c));
    };
    dist[s] = 0.0;
    sigma[s] = 1.0;
    push_or_decrease(s, 0.0);
    while (hlen > 0) {
        double d = hk[0];
        uint32_t v = hv[0];
        pos[v] = kNone;
        hlen--;
        if (hlen > 0) {
            hk[0] = hk[hlen];
            hv[0] = hv[hlen];
            sift_down(0);
        }
        rank[v] = (uint32_t)norder;
        order[norder++] = v;
        for (uint32_t a = off[v]; a < off[v + 1]; a++) {
            uint32_t w = idx[a];
            if (rank[w] != kNone) continue;
            double alt = d + lens[a];
            double cur = dist[w];
            if (isinf(cur)) {
                dist[w] = alt;
                sigma[w] = sigma[v];
                push_or_decrease(w, alt);
            } else if (approx_eq(alt, cur)) {
                sigma[w] += sigma[v];
            } else if (alt < cur) {
                dist[w] = alt;
                sigma[w] = sigma[v];
                push_or_decrease(w, alt);
            }
        }
    }
    double mx = 0.0;
    for (size_t i = norder; i-- > 0;) {
        uint32_t v = order[i];
        uint32_t rv = rank[v];
        double dvv = dist[v];
        double dv = 0.0;
        for (uint32_t a = off[v]; a < off[v + 1]; a++) {
            uint32_t w = idx[a];
            if (rank[w] > rv && rank[w] != kNone && approx_eq(dvv + lens[a], dist[w]))
                dv += coef[w];
        }
        double delta = sigma[v] * dv;
        coef[v] = (1.0 + delta) / sigma[v];
        size_t mi = (size_t)b * n + v;
        atomicOr((unsigned long long*)&vis[mi], 1ull << l);
        sig[mi * kLanes + l] = delta;
        mx = fmax(mx, sigma[v]);
    }
    atomicMax(&ctrl->max_sigma_bits, (unsigned long long)__double_as_longlong(mx));
    if (mx > kTwo53) ctrl->sig_overflow = 1;
}

// ---- cross-source reduction -----------------------------------------------

// Adds this super-batch's dependencies into the per-node group
// accumulator in item order. When an item belongs to a later group than
// the last one added, the accumulator is first merged into bc and reset
// (groups without contributions would add +0.0, which changes nothing).
// This is sum_in_order's sequence of operations on every node.
__global__ void accumulate_kernel(const uint64_t* __restrict__ vis, const double* __restrict__ sig,
                                  int slots, size_t n, int64_t item0, int64_t k, int64_t gsize,
                                  const int64_t* __restrict__ list, double* __restrict__ acc,
                                  double* __restrict__ bc, int64_t* __restrict__ gcur,
                                  const uint32_t* __restrict__ off, Ctrl* ctrl) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long rn = 0, ra = 0;
    if (v < n) {
        double a = acc[v], B = bc[v];
        int64_t g = gcur[v];
        for (int b = 0; b < slots; b++) {
            int64_t ib = item0 + (int64_t)b * kLanes;
            if (ib >= k) break;
            uint64_t M = vis[(size_t)b * n + v];
            rn += (unsigned long long)__popcll(M);
            while (M) {
                int l = __ffsll((unsigned long long)M) - 1;
                M &= M - 1;
                int64_t i = ib + l;
                int64_t s = list ? list[i] : i;
                if ((size_t)s == v) continue;
                int64_t gi = i / gsize;
                if (gi != g) {
                    B += a;
                    a = 0.0;
                    g = gi;
                }
                a += sig[((size_t)b * n + v) * kLanes + l];
            }
        }
        acc[v] = a;
        bc[v] = B;
        gcur[v] = g;
        ra = rn * (unsigned long long)(off[v + 1] - off[v]);
    }
    rn = wave_sum_u64(rn);
    ra = wave_sum_u64(ra);
    if (lane_id() == 0 && (rn || ra)) {
        atomicAdd(&ctrl->reached_nodes, rn);
        atomicAdd(&ctrl->reached_arcs, ra);
    }
}

// Final merge of the last group.
__global__ void finish_kernel(double* __restrict__ bc, const double* __restrict__ acc, size_t n) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) bc[v] += acc[v];
}

__global__ void scale_kernel(double* __restrict__ bc, size_t n, double scale) {
    size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) bc[v] *= scale;
}

inline unsigned blocks_for_threads(size_t threads) {
    size_t b = (threads + kBlock - 1) / kBlock;
    return (unsigned)std::max<size_t>(1, std::min<size_t>(b, 1u << 20));
}
inline unsigned blocks_for_waves(size_t waves) { return blocks_for_threads(waves * 64); }

// Device scratch of the batched engine, cached on the graph handle.
struct Workspace {
    int slots = 0;
    uint64_t n = 0;
    DevBuf<uint64_t> vis, cand, g0, g1, inb;
    DevBuf<double> sig, coef64, acc;
    DevBuf<float> coef32;
    DevBuf<uint32_t> dist, blen;
    DevBuf<int64_t> gcur, list;
    DevBuf<QEnt> q;
    size_t qcap = 0;
    DevBuf<uint2> touched, blists;
    size_t list_cap = 0;
    DevBuf<Ctrl> ctrl;
    Ctrl* host_ctrl = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evd0 = nullptr, evd1 = nullptr;
    ~Workspace() {
        if (host_ctrl) (void)hipHostFree(host_ctrl);
        for (hipEvent_t e : {ev0, ev1, evd0, evd1})
            if (e) (void)hipEventDestroy(e);
    }
};

void free_workspace(void* p) { delete static_cast<Workspace*>(p); }

Workspace& get_workspace(Graph& g) {
    if (!g.workspace) {
        auto* ws = new Workspace();
        g.workspace = ws;
        g.workspace_free = free_workspace;
        NETBC_HIP(hipHostMalloc((void**)&ws->host_ctrl, sizeof(Ctrl), hipHostMallocDefault));
        NETBC_HIP(hipEventCreate(&ws->ev0));
        NETBC_HIP(hipEventCreate(&ws->ev1));
        NETBC_HIP(hipEventCreate(&ws->evd0));
        NETBC_HIP(hipEventCreate(&ws->evd1));
    }
    return *static_cast<Workspace*>(g.workspace);
}

void read_ctrl(Workspace& ws, hipStream_t st) {
    NETBC_HIP(hipMemcpyAsync(ws.host_ctrl, ws.ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
    NETBC_HIP(hipStreamSynchronize(st));
}

// Sizes (in bytes per node and slot) of the per-slot arrays.
size_t slot_bytes_per_node(const Lengths& L, int precision) {
    size_t b = 3 * 8 + kLanes * 8 + kLanes * (precision ? 4 : 8);
    if (L.kind == kInteger) b += kLanes * 4 + 8 * ((L.max + 1 + 63) / 64) + 8 * (L.max + 1);
    if (L.kind == kReal) b += 0;
    return b;
}

int choose_slots(const Graph& g, const Lengths& L, int64_t k, const RunOptions& opt) {
    int64_t batches = (k + kLanes - 1) / kLanes;
    int want = opt.slots > 0 ? std::min(opt.slots, 1024) : 16;
    if (L.kind == kReal) want = opt.slots > 0 ? opt.slots : 4;
    size_t per = slot_bytes_per_node(L, opt.precision) * std::max<uint64_t>(g.n, 1);
    size_t budget = (size_t)48 << 30;
    int by_mem = (int)std::max<size_t>(1, budget / std::max<size_t>(per, 1));
    int s = std::min<int64_t>({(int64_t)want, batches, (int64_t)by_mem});
    return std::max(1, s);
}

}  // namespace

void run_batched(Graph& g, const Lengths& L, const int64_t* sources, int64_t k,
                 const RunOptions& opt, double* d_bc, RunDiag& diag);

namespace {

// One super-batch: forward search and backward sweep for `slots` batches
// starting at item0. Returns false if the level queue overflowed (the
// caller enlarges it and repeats).
struct Engine {
    Graph& g;
    const Lengths& L;
    Workspace& ws;
    const RunOptions& opt;
    RunDiag& diag;
    hipStream_t st;
    size_t n;
    int slots;
    const int64_t* d_list;
    int64_t k;

    // Per-level boundaries in the queue (level d = [lv[d], lv[d+1])).
    std::vector<uint32_t> lv;
    std::vector<uint32_t> lv_dist;  // Dial: distance of each level

    float dom_ms = 0.0f;

    void launch_check() { NETBC_HIP(hipGetLastError()); }

    void reset_slots() {
        size_t nm = (size_t)slots * n;
        unsigned gb = blocks_for_threads(std::min<size_t>(nm, (size_t)1 << 22));
        fill_u64_kernel<<<gb, kBlock, 0, st>>>(ws.vis.p, nm, 0);
        launch_check();
        if (L.kind == kUniform) {
            fill_u64_kernel<<<gb, kBlock, 0, st>>>(ws.cand.p, nm, 0);
            fill_u64_kernel<<<gb, kBlock, 0, st>>>(ws.g0.p, nm, 0);
            fill_u64_kernel<<<gb, kBlock, 0, st>>>(ws.g1.p, nm, 0);
            launch_check();
        }
        if (L.kind == kInteger) {
            fill_u32_kernel<<<blocks_for_threads(std::min<size_t>(nm * kLanes, (size_t)1 << 22)),
                              kBlock, 0, st>>>(ws.dist.p, nm * kLanes, kNone);
            int words = (int)((L.max + 1 + 63) / 64);
            fill_u64_kernel<<<gb, kBlock, 0, st>>>(ws.inb.p, nm * words, 0);
            fill_u32_kernel<<<1, kBlock, 0, st>>>(ws.blen.p, L.max + 1, 0);
            launch_check();
        }
        NETBC_HIP(hipMemsetAsync(ws.ctrl.p, 0, offsetof(Ctrl, sig_overflow), st));
    }

    template <class C>
    C* coef_ptr();

    bool run_bfs(int64_t item0) {
        const DeviceGraph& d = g.dev;
        reset_slots();
        init_sources_kernel<<<blocks_for_threads((size_t)slots * kLanes), kBlock, 0, st>>>(
            d_list, item0, k, slots, n, ws.vis.p, ws.g0.p, ws.sig.p, ws.q.p, ws.ctrl.p, nullptr);
        launch_check();
        read_ctrl(ws, st);
        lv.assign(1, 0);
        lv.push_back(ws.host_ctrl->qlen);
        uint64_t* G[2] = {ws.g0.p, ws.g1.p};
        unsigned sigma_blocks = blocks_for_waves(std::min<size_t>((size_t)slots * n, 1u << 16));
        for (int dl = 0;; dl++) {
            uint32_t qb = lv[dl], qe = lv[dl + 1];
            uint32_t zb = dl > 0 ? lv[dl - 1] : 0, ze = dl > 0 ? lv[dl] : 0;
            size_t waves = std::max<size_t>(qe - qb, (ze - zb + 63) / 64);
            bfs_expand_kernel<<<blocks_for_waves(waves), kBlock, 0, st>>>(
                ws.q.p, qb, qe, zb, ze, d.out_off.p, d.out_idx.p, ws.vis.p, ws.cand.p,
                G[(dl + 1) & 1], ws.touched.p, &ws.ctrl.p->tlen[dl & 1], n);
            launch_check();
            bfs_sigma_kernel<<<sigma_blocks, kBlock, 0, st>>>(
                ws.touched.p, &ws.ctrl.p->tlen[dl & 1], &ws.ctrl.p->tlen[(dl + 1) & 1],
                d.in_off_p(), d.in_idx_p(), G[dl & 1], ws.cand.p, ws.vis.p, G[(dl + 1) & 1],
                ws.sig.p, ws.q.p, (uint32_t)ws.qcap, ws.ctrl.p, n);
            launch_check();
            read_ctrl(ws, st);
            if (ws.host_ctrl->qoverflow) return false;
            uint32_t ql = ws.host_ctrl->qlen;
            if (ql == qe) break;
            lv.push_back(ql);
            diag.levels++;
        }
        return true;
    }

    template <class C>
    void backward_bfs(C* coef) {
        const DeviceGraph& d = g.dev;
        uint64_t* G[2] = {ws.g0.p, ws.g1.p};
        int D = (int)lv.size() - 2;  // last level with entries
        for (int dl = D; dl >= 1; dl--) {
            uint32_t qb = lv[dl], qe = lv[dl + 1];
            uint32_t ab = dl + 2 <= D ? lv[dl + 2] : 0, ae = dl + 2 <= D ? lv[dl + 3] : 0;
            size_t waves = std::max<size_t>(qe - qb, (ae - ab + 63) / 64);
            NETBC_HIP(hipEventRecord(ws.evd0, st));
            bfs_backward_kernel<C><<<blocks_for_waves(waves), kBlock, 0, st>>>(
                ws.q.p, qb, qe, ab, ae, dl == D, d.out_off.p, d.out_idx.p, G[(dl + 1) & 1],
                G[dl & 1], ws.sig.p, coef, n);
            launch_check();
            NETBC_HIP(hipEventRecord(ws.evd1, st));
            NETBC_HIP(hipEventSynchronize(ws.evd1));
            float ms = 0.0f;
            NETBC_HIP(hipEventElapsedTime(&ms, ws.evd0, ws.evd1));
            dom_ms += ms;
            diag.dominant_launches++;
        }
    }

    bool run_dial(int64_t item0) {
        const DeviceGraph& d = g.dev;
        reset_slots();
        init_sources_kernel<<<blocks_for_threads((size_t)slots * kLanes), kBlock, 0, st>>>(
            d_list, item0, k, slots, n, ws.vis.p, nullptr, ws.sig.p, ws.q.p, ws.ctrl.p,
            ws.dist.p);
        launch_check();
        read_ctrl(ws, st);
        lv.assign(1, 0);
        lv.push_back(ws.host_ctrl->qlen);
        lv_dist.assign(1, 0);
        int W = (int)L.max + 1;
        int words = (W + 63) / 64;
        unsigned settle_blocks = blocks_for_waves(std::min<size_t>((size_t)slots * n, 1u << 16));
        const uint32_t* in_len = d.directed ? L.in_len.p : L.out_len.p;
        // Distance 0: relax the sources' arcs.
        uint32_t D = 0;
        for (;;) {
            uint32_t qb = lv[lv.size() - 2], qe = lv.back();
            bool have = lv_dist.back() == D && qe > qb;
            dial_relax_kernel<<<blocks_for_waves(have ? qe - qb : 1), kBlock, 0, st>>>(
                ws.q.p, have ? qb : 0, have ? qe : 0, D, W, words, (int)(D % (uint32_t)W),
                ws.blen.p, ws.blists.p, ws.list_cap, ws.inb.p, d.out_off.p, d.out_idx.p,
                L.out_len.p, ws.dist.p, ws.vis.p, ws.ctrl.p, n);
            launch_check();
            read_ctrl(ws, st);
            if (ws.host_ctrl->qoverflow) return false;
            if (ws.host_ctrl->pending == 0) break;
            D++;
            int kb = (int)(D % (uint32_t)W);
            dial_settle_kernel<<<settle_blocks, kBlock, 0, st>>>(
                ws.blists.p + (size_t)kb * ws.list_cap, ws.blen.p, ws.list_cap, D, kb, words,
                ws.inb.p,
                d.in_off_p(), d.in_idx_p(), in_len, ws.dist.p, ws.vis.p, ws.sig.p, ws.q.p,
                (uint32_t)ws.qcap, ws.ctrl.p, n);
            launch_check();
            read_ctrl(ws, st);
            if (ws.host_ctrl->qoverflow) return false;
            uint32_t ql = ws.host_ctrl->qlen;
            if (ql > lv.back()) {
                lv.push_back(ql);
                lv_dist.push_back(D);
                diag.levels++;
            }
        }
        return true;
    }

    template <class C>
    void backward_dial(C* coef) {
        const DeviceGraph& d = g.dev;
        int D = (int)lv.size() - 2;
        for (int dl = D; dl >= 1; dl--) {
            uint32_t qb = lv[dl], qe = lv[dl + 1];
            NETBC_HIP(hipEventRecord(ws.evd0, st));
            dial_backward_kernel<C><<<blocks_for_waves(qe - qb), kBlock, 0, st>>>(
                ws.q.p, qb, qe, lv_dist[dl], d.out_off.p, d.out_idx.p, L.out_len.p, ws.dist.p,
                ws.vis.p, ws.sig.p, coef, n);
            launch_check();
            NETBC_HIP(hipEventRecord(ws.evd1, st));
            NETBC_HIP(hipEventSynchronize(ws.evd1));
            float ms = 0.0f;
            NETBC_HIP(hipEventElapsedTime(&ms, ws.evd0, ws.evd1));
            dom_ms += ms;
            diag.dominant_launches++;
        }
    }

    void run_dijkstra(int64_t item0, DjkState& dst) {
        const DeviceGraph& d = g.dev;
        reset_slots();
        NETBC_HIP(hipEventRecord(ws.evd0, st));
        dijkstra_kernel<<<slots, kLanes, 0, st>>>(
            d_list, item0, k, slots, n, d.out_off.p, d.out_idx.p, L.out_real.p, dst, ws.vis.p,
            ws.sig.p, ws.ctrl.p);
        launch_check();
        NETBC_HIP(hipEventRecord(ws.evd1, st));
        NETBC_HIP(hipEventSynchronize(ws.evd1));
        float ms = 0.0f;
        NETBC_HIP(hipEventElapsedTime(&ms, ws.evd0, ws.evd1));
        dom_ms += ms;
        diag.dominant_launches++;
    }
};

// Level-queue entries: every source entry plus four levels per node and
// slot to start with; the queue grows if a run needs more.
size_t queue_capacity(int slots, size_t n) {
    return std::min<size_t>((size_t)slots * (n * 4 + kLanes) + 1024, 0xFFFFFFFFull);
}

void grow_queue(Workspace& ws) {
    if (ws.qcap >= 0xFFFFFFFFull) throw Error("level queue exceeds 2^32 entries");
    size_t nq = std::min<size_t>(ws.qcap * 2, 0xFFFFFFFFull);
    ws.q.release();
    ws.q.alloc(nq);
    ws.qcap = nq;
}

void ensure_workspace(Workspace& ws, const Graph& g, const Lengths& L, int slots, int precision,
                      size_t qcap) {
    size_t n = std::max<uint64_t>(g.n, 1);
    size_t nm = (size_t)slots * n;
    ws.slots = slots;
    ws.n = g.n;
    ws.vis.alloc(nm);
    ws.sig.alloc(nm * kLanes);
    if (L.kind == kUniform) {
        ws.cand.alloc(nm);
        ws.g0.alloc(nm);
        ws.g1.alloc(nm);
        ws.touched.alloc(nm);
    }
    if (L.kind != kReal) {
        if (precision) ws.coef32.alloc(nm * kLanes);
        else ws.coef64.alloc(nm * kLanes);
        if (qcap > ws.qcap || !ws.q.p) {
            ws.q.release();
            ws.q.alloc(qcap);
            ws.qcap = qcap;
        }
    }
    if (L.kind == kInteger) {
        int W = (int)L.max + 1;
        int words = (W + 63) / 64;
        ws.dist.alloc(nm * kLanes);
        ws.inb.alloc(nm * words);
        ws.blen.alloc(W);
        if (ws.list_cap < nm || !ws.blists.p) {
            ws.blists.release();
            ws.blists.alloc((size_t)W * nm);
            ws.list_cap = nm;
        }
    }
    ws.acc.alloc(n);
    ws.gcur.alloc(n);
    ws.ctrl.alloc(1);
}

}  // namespace

void run_batched(Graph& g, const Lengths& L, const int64_t* sources, int64_t k,
                 const RunOptions& opt, double* d_bc, RunDiag& diag) {
    hipStream_t st = g.stream;
    size_t n = g.n;
    diag.path_kind = L.kind;
    if (n == 0) return;
    NETBC_HIP(hipMemsetAsync(d_bc, 0, n * sizeof(double), st));
    if (k == 0) return;
    Workspace& ws = get_workspace(g);
    int slots = choose_slots(g, L, k, opt);
    diag.slots = slots;
    size_t qcap = std::max<size_t>(queue_capacity(slots, n), ws.qcap);
    ensure_workspace(ws, g, L, slots, opt.precision, qcap);
    NETBC_HIP(hipMemsetAsync(ws.acc.p, 0, n * sizeof(double), st));
    NETBC_HIP(hipMemsetAsync(ws.gcur.p, 0, n * sizeof(int64_t), st));
    const int64_t* d_list = nullptr;
    if (sources) {
        ws.list.alloc(k);
        NETBC_HIP(hipMemcpyAsync(ws.list.p, sources, k * sizeof(int64_t), hipMemcpyHostToDevice,
                                 st));
        d_list = ws.list.p;
    }
    NETBC_HIP(hipMemsetAsync(ws.ctrl.p, 0, sizeof(Ctrl), st));
    DevBuf<double> djk_buf;
    DevBuf<uint32_t> djk_u32;
    DjkState djk{};
    if (L.kind == kReal) {
        size_t T = (size_t)slots * kLanes * n;
        djk_buf.alloc(4 * T);
        djk_u32.alloc(4 * T);
        djk = DjkState{djk_buf.p, djk_buf.p + T, djk_buf.p + 2 * T, djk_u32.p, djk_u32.p + T,
                       djk_buf.p + 3 * T, djk_u32.p + 2 * T, djk_u32.p + 3 * T};
    }
    Engine eng{g, L, ws, opt, diag, st, n, slots, d_list, k};
    int64_t gsize = (k + 63) / 64;
    NETBC_HIP(hipEventRecord(ws.ev0, st));
    for (int64_t item0 = 0; item0 < k; item0 += (int64_t)slots * kLanes) {
        if (L.kind == kReal) {
            eng.run_dijkstra(item0, djk);
        } else {
            for (;;) {
                bool ok = L.kind == kUniform ? eng.run_bfs(item0) : eng.run_dial(item0);
                if (ok) break;
                // Level queue overflow: grow and repeat the super-batch.
                if (ws.host_ctrl->qoverflow == 2) throw Error("bucket list overflow");
                grow_queue(ws);
            }
            if (opt.precision) {
                if (L.kind == kUniform) eng.backward_bfs<float>(ws.coef32.p);
                else eng.backward_dial<float>(ws.coef32.p);
            } else {
                if (L.kind == kUniform) eng.backward_bfs<double>(ws.coef64.p);
                else eng.backward_dial<double>(ws.coef64.p);
            }
        }
        accumulate_kernel<<<blocks_for_threads(n), kBlock, 0, st>>>(
            ws.vis.p, ws.sig.p, slots, n, item0, k, gsize, d_list, ws.acc.p, d_bc, ws.gcur.p,
            g.dev.out_off.p, ws.ctrl.p);
        NETBC_HIP(hipGetLastError());
    }
    finish_kernel<<<blocks_for_threads(n), kBlock, 0, st>>>(d_bc, ws.acc.p, n);
    NETBC_HIP(hipGetLastError());
    NETBC_HIP(hipEventRecord(ws.ev1, st));
    read_ctrl(ws, st);
    float ms = 0.0f;
    NETBC_HIP(hipEventElapsedTime(&ms, ws.ev0, ws.ev1));
    diag.kernel_ms += ms;
    diag.dominant_ms += eng.dom_ms;
    Ctrl& c = *ws.host_ctrl;
    double mx;
    unsigned long long bits = c.max_sigma_bits;
    memcpy(&mx, &bits, 8);
    diag.max_sigma = std::max(diag.max_sigma, mx);
    diag.sigma_overflow |= c.sig_overflow ? 1 : 0;
    diag.reached_nodes += (int64_t)c.reached_nodes;
    diag.reached_arcs += (int64_t)c.reached_arcs;
}

void finish_scores(Graph& g, bool normalized, double* d_bc, double* out) {
    hipStream_t st = g.stream;
    uint64_t n = g.n;
    if (n == 0) return;
    // betweenness.rs:100-113
    double scale = g.directed ? 1.0 : 0.5;
    if (normalized && n > 2) {
        double pairs = (double)((n - 1) * (n - 2));
        scale *= g.directed ? 1.0 / pairs : 2.0 / pairs;
    }
    if (scale != 1.0) {
        scale_kernel<<<blocks_for_threads(n), kBlock, 0, st>>>(d_bc, n, scale);
        NETBC_HIP(hipGetLastError());
    }
    NETBC_HIP(hipMemcpyAsync(out, d_bc, n * sizeof(double), hipMemcpyDeviceToHost, st));
    NETBC_HIP(hipStreamSynchronize(st));
}

void run_dump(Graph& g, const Lengths& L, const int64_t* sources, int64_t k,
              const RunOptions& opt, uint64_t* dist, double* sigma, double* delta) {
    // Runs the batched engine one super-batch at a time and copies each
    // source's state out after the forward search (levels, sigma) and after
    // the backward sweep (delta).
    hipStream_t st = g.stream;
    size_t n = g.n;
    if (n == 0 || k == 0) return;
    if (L.kind == kReal) throw Error("the state dump covers breadth-first and Dial searches");
    Workspace& ws = get_workspace(g);
    RunOptions o = opt;
    int slots = choose_slots(g, L, k, o);
    size_t qcap = std::max<size_t>(queue_capacity(slots, n), ws.qcap);
    ensure_workspace(ws, g, L, slots, o.precision, qcap);
    ws.list.alloc(k);
    NETBC_HIP(hipMemcpyAsync(ws.list.p, sources, k * sizeof(int64_t), hipMemcpyHostToDevice, st));
    NETBC_HIP(hipMemsetAsync(ws.ctrl.p, 0, sizeof(Ctrl), st));
    RunDiag diag;
    Engine eng{g, L, ws, o, diag, st, n, slots, ws.list.p, k};
    std::vector<double> rows((size_t)slots * n * kLanes);
    std::vector<uint32_t> drows;
    std::vector<uint64_t> vis((size_t)slots * n);
    for (int64_t item0 = 0; item0 < k; item0 += (int64_t)slots * kLanes) {
        for (;;) {
            bool ok = L.kind == kUniform ? eng.run_bfs(item0) : eng.run_dial(item0);
            if (ok) break;
            if (ws.host_ctrl->qoverflow == 2) throw Error("bucket list overflow");
            grow_queue(ws);
        }
        int64_t cnt = std::min<int64_t>((int64_t)slots * kLanes, k - item0);
        NETBC_HIP(hipMemcpyAsync(rows.data(), ws.sig.p, rows.size() * 8, hipMemcpyDeviceToHost, st));
        NETBC_HIP(hipMemcpyAsync(vis.data(), ws.vis.p, vis.size() * 8, hipMemcpyDeviceToHost, st));
        if (L.kind == kInteger) {
            drows.resize((size_t)slots * n * kLanes);
            NETBC_HIP(hipMemcpyAsync(drows.data(), ws.dist.p, drows.size() * 4,
                                     hipMemcpyDeviceToHost, st));
        }
        std::vector<QEnt> q(eng.lv.back());
        if (!q.empty())
            NETBC_HIP(hipMemcpyAsync(q.data(), ws.q.p, q.size() * sizeof(QEnt),
                                     hipMemcpyDeviceToHost, st));
        NETBC_HIP(hipStreamSynchronize(st));
        for (int64_t t = 0; t < cnt; t++) {
            int b = (int)(t / kLanes), l = (int)(t % kLanes);
            size_t o2 = (size_t)(item0 + t) * n;
            for (size_t v = 0; v < n; v++) {
                bool r = (vis[(size_t)b * n + v] >> l) & 1;
                sigma[o2 + v] = r ? rows[((size_t)b * n + v) * kLanes + l] : 0.0;
                dist[o2 + v] = UINT64_MAX;
                if (r && L.kind == kInteger) dist[o2 + v] = drows[((size_t)b * n + v) * kLanes + l];
                delta[o2 + v] = 0.0;
            }
        }
        if (L.kind == kUniform) {
            for (size_t dl = 0; dl + 1 < eng.lv.size(); dl++)
                for (uint32_t i = eng.lv[dl]; i < eng.lv[dl + 1]; i++) {
                    const QEnt& e = q[i];
                    uint64_t m = e.m;
                    while (m) {
                        int l = __builtin_ctzll(m);
                        m &= m - 1;
                        int64_t t = (int64_t)e.b * kLanes + l;
                        if (t < cnt) dist[(size_t)(item0 + t) * n + e.v] = dl;
                    }
                }
        }
        if (o.precision) {
            if (L.kind == kUniform) eng.backward_bfs<float>(ws.coef32.p);
            else eng.backward_dial<float>(ws.coef32.p);
        } else {
            if (L.kind == kUniform) eng.backward_bfs<double>(ws.coef64.p);
            else eng.backward_dial<double>(ws.coef64.p);
        }
        NETBC_HIP(hipMemcpyAsync(rows.data(), ws.sig.p, rows.size() * 8, hipMemcpyDeviceToHost, st));
        NETBC_HIP(hipStreamSynchronize(st));
        for (int64_t t = 0; t < cnt; t++) {
            int b = (int)(t / kLanes), l = (int)(t % kLanes);
            size_t o2 = (size_t)(item0 + t) * n;
            int64_t s = sources[item0 + t];
            for (size_t v = 0; v < n; v++) {
                bool r = (vis[(size_t)b * n + v] >> l) & 1;
                if (r && (int64_t)v != s) delta[o2 + v] = rows[((size_t)b * n + v) * kLanes + l];
            }
        }
    }
}

}  // namespace netbc
