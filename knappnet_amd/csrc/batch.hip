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
#include <map>
#include <mutex>
#include <type_traits>
#include <unordered_map>

#include "netbc_internal.h"
#include "tuning.h"

namespace netbc {
namespace {

struct QEnt {
    uint32_t v;
    uint32_t b;
    unsigned long long m;
};

// Per-super-batch counters come first (cleared before every super-batch);
// the rest accumulate over the whole run.
// Counters summed by the host are spread over this many addresses
// (blockIdx.x % kSpread picks one).
constexpr int kSpread = 32;

struct Ctrl {
    uint32_t qlen;
    uint32_t tlen[2];
    uint32_t qoverflow;
    uint32_t hlen;       // heavy-row entries listed (positions in the level queue)
    uint32_t left[2];    // bucketed searches: tentative distances still pending
    uint32_t hchunks;    // chunk tasks of the heavy rows listed last
    // Bucketed searches: entries with long rows, listed from the end of the
    // bucket list (alternating by step parity like tlen).
    uint32_t tlong[2];
    // Direction choice: out-arcs of the entries of a level (alternating by
    // level parity) and in-arcs of the (slot, node) pairs whose active lanes
    // are all reached.
    // Both are spread over kSpread counters (summed by the host) so that
    // the wavefronts of a launch do not all add into one address.
    unsigned long long front_arcs[2][32];
    unsigned long long closed_arcs[32];
    // Bucketed searches: the lowest bucket still pending after a step
    // (alternating by step parity; all ones when none).
    unsigned long long next_bucket[2];
    // Integer lengths (distance ring): the largest distance offered so far.
    uint32_t last_offer;
    uint32_t pad1;
    // Work tickets of the dependency sweep, one counter pair per class of
    // workgroups (blockIdx.x % 32, four classes per XCD), each in its own
    // 64-byte line; the last wavefront of a class to finish resets its pair
    // for the next launch.
    struct XTicket {
        uint32_t ticket, exits;
        uint32_t pad[14];
    } xticket[32];
    uint32_t sig_overflow;
    uint32_t pad2;
    unsigned long long max_sigma_bits;
    unsigned long long reached_nodes[32];  // spread as front_arcs
    unsigned long long reached_arcs[32];
    unsigned long long fwd_arcs;  // arcs scanned by top-down expansion
    // Bytes of the reachability-only bottom-up steps (bfs_pull_kernel
    // without path counts), spread as front_arcs: 8 per pair examined, 8 per
    // in-row looked up plus 12 per in-arc read, 40 per pair reached (visited
    // and level words, out-row bounds, queue entry).
    unsigned long long reach_bytes[32];
    // Work counters (filled only in builds with NETBC_COUNTERS):
    // 0 expand entries, 1 expand arcs, 2 arcs with offers, 3 offered lanes
    // (integer lengths: 0 settled pairs, 1 their in-arcs, 2 out-arcs, 3 offers),
    // 4 sweep entries, 5 sweep arcs, 6 sweep contributing arcs, 7 their lanes,
    // 8 bottom-up arcs, 9 their hits, 10 gathered arcs, 11 their hits,
    // 12 hub-task arcs, 13 their hits, 14 hub tasks with lanes sought,
    // 15 bottom-up pairs with lanes sought.
    unsigned long long work[16];
};

#ifdef NETBC_COUNTERS
#define NETBC_COUNT(var, x) (var) += (x)
#define NETBC_FLUSH(ctrl, i, var)                                             \
    do {                                                                      \
        unsigned long long t_ = wave_sum_u64(var);                            \
        if (lane_id() == 0 && t_) atomicAdd(&(ctrl)->work[i], t_);            \
    } while (0)
#else
#define NETBC_COUNT(var, x) (void)0
#define NETBC_FLUSH(ctrl, i, var) (void)0
#endif

constexpr double kTwo53 = 9007199254740992.0;
// Rows loaded ahead of the ordered sum in the dependency sweeps.
#ifndef NETBC_AHEAD
#define NETBC_AHEAD 8
#endif
constexpr int kAhead = NETBC_AHEAD;
// The same for the bucketed sweep's unpacked walk (dial_backward_kernel,
// entries wider than kPackLanes), whose registers are capped for 7
// wavefronts per SIMD. Measured on BA n=200,000 with lengths 1-16 (16,384
// sources): 4 hits per batch at 7 wavefronts per SIMD 443 -> 373 ms against
// 8 hits at 5 (its register use).
#ifndef NETBC_DIAL_BACK_AHEAD
#define NETBC_DIAL_BACK_AHEAD 4
#endif
constexpr int kDialBackAhead = NETBC_DIAL_BACK_AHEAD;
#ifndef NETBC_DIAL_GROUP
#define NETBC_DIAL_GROUP 2
#endif
// Entries per work ticket in the Dial step on long bucket lists (one per
// ticket measured 5% slower on BA n = 200,000, four 6% slower); the sweep
// takes one. A power of two.
constexpr uint32_t kDialGroup = NETBC_DIAL_GROUP;
#ifndef NETBC_DIAL_BACK_WAVES
#define NETBC_DIAL_BACK_WAVES 7
#endif
// Predecessor rows loaded at once by the path-count gathers (gather_counts).
#ifndef NETBC_GATHER_AHEAD
#define NETBC_GATHER_AHEAD 8
#endif
constexpr int kGatherAhead = NETBC_GATHER_AHEAD;
// Level bitmap buffers per slot: levels below kKeptLevels keep their own
// buffer from the forward search to the sweep, deeper ones alternate
// between the last two (then the sweep rebuilds them from the queue).
constexpr int kLevelBufs = 16;
constexpr int kKeptLevels = kLevelBufs - 2;
// Touched nodes staged per wave before one append to the shared list.
constexpr uint32_t kTouchBuf = 256;
// Queue entries staged per wavefront by the bottom-up step (16 bytes each).
constexpr uint32_t kPullBuf = 192;
constexpr size_t kPullLds = (size_t)(kBlock / 64) * kPullBuf * 16;
// A level is searched bottom-up when the out-arcs of its entries times
// this factor exceed the in-arcs of the pairs with unreached lanes.
#ifndef NETBC_ALPHA
#define NETBC_ALPHA 1.0
#endif
constexpr double kBottomUpAlpha = NETBC_ALPHA;

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
__device__ __forceinline__ uint32_t rl(uint32_t x, int j) { return rl32(x, j); }
__device__ __forceinline__ double rl(double x, int j) {
    return __longlong_as_double((long long)rl64((uint64_t)__double_as_longlong(x), j));
}
// 1 for a target node (target == nullptr: every node), else 0.
__device__ __forceinline__ double target_of(const uint8_t* __restrict__ target, uint32_t v) {
    return target ? (double)target[v] : 1.0;
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

// Publishes a wave's largest path count (one atomic per wave, at the end of
// a kernel) and flags counts beyond 2^53.
__device__ __forceinline__ void note_sigma(Ctrl* ctrl, double s) {
    double mx = wave_max(s);
    if (lane_id() == 0 && mx > 0.0) {
        unsigned long long bits = (unsigned long long)__double_as_longlong(mx);
        // a plain read first: most wavefronts hold no new maximum (a stale
        // value only lets an atomic through)
        if (bits > *(volatile unsigned long long*)&ctrl->max_sigma_bits)
            atomicMax(&ctrl->max_sigma_bits, bits);
        if (mx > kTwo53) ctrl->sig_overflow = 1;
    }
}

// ---- initialisation -------------------------------------------------------

// Marks the source of every active lane: visited, sigma = 1, level 0
// (distance 0 in the bucketed searches).
template <class Dist>
__global__ void init_sources_kernel(const int64_t* __restrict__ list, int64_t item0, int64_t k,
                                    int slots, size_t n, uint64_t* __restrict__ vis,
                                    uint64_t* __restrict__ g0, double* __restrict__ sig,
                                    Dist* __restrict__ dist, uint64_t* __restrict__ pend,
                                    int words) {
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
    if (dist) dist[mi * kLanes + l] = Dist(0);
    if (pend) atomicOr((unsigned long long*)&pend[mi * words], 1ull);  // bucket of distance 0
}

// The level-0 queue entries: one per (slot, source node), holding every
// lane of the slot that starts at that node (a source listed twice in one
// slot gets one entry with two lanes, as every later level has), written
// by the lowest such lane.
__global__ void source_entries_kernel(const int64_t* __restrict__ list, int64_t item0, int64_t k,
                                      int slots, size_t n, const uint64_t* __restrict__ vis,
                                      QEnt* __restrict__ q, Ctrl* ctrl,
                                      const uint8_t* __restrict__ heavy_flag,
                                      uint32_t* __restrict__ hq,
                                      const uint32_t* __restrict__ out_off) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= slots * kLanes) return;
    int b = t / kLanes, l = t % kLanes;
    int64_t i = item0 + t;
    if (i >= k) return;
    uint32_t s = (uint32_t)(list ? list[i] : i);
    const uint64_t M = vis[(size_t)b * n + s];
    if (__ffsll((unsigned long long)M) - 1 != l) return;
    uint32_t pos = atomicAdd(&ctrl->qlen, 1u);
    q[pos] = QEnt{s, (uint32_t)b, (unsigned long long)M};
    if (hq && heavy_flag[s]) hq[atomicAdd(&ctrl->hlen, 1u)] = pos;
    if (out_off)
        atomicAdd(&ctrl->front_arcs[0][s % kSpread], (unsigned long long)(out_off[s + 1] - out_off[s]));
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, size_t count, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = v;
}

// Zeroes three buffers of `count` words each in one pass, 16 bytes per
// store (the buffers are 256-byte aligned): one launch instead of three.
__global__ void zero3_kernel(uint64_t* __restrict__ a, uint64_t* __restrict__ b,
                             uint64_t* __restrict__ c, size_t count) {
    const size_t pairs = count / 2;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const ulonglong2 z = make_ulonglong2(0, 0);
    for (size_t i = tid; i < pairs; i += stride) {
        reinterpret_cast<ulonglong2*>(a)[i] = z;
        reinterpret_cast<ulonglong2*>(b)[i] = z;
        reinterpret_cast<ulonglong2*>(c)[i] = z;
    }
    if ((count & 1) && tid == 0) {
        a[count - 1] = 0;
        b[count - 1] = 0;
        c[count - 1] = 0;
    }
}

__global__ void fill_u64_kernel(uint64_t* __restrict__ p, size_t count, uint64_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = v;
}

// ---- breadth-first search: forward -----------------------------------------

// Chunk counts of the heavy rows listed in hq[hbeg, *hend): hpre[i] = first
// chunk task of entry i (relative to hbeg); *total = all of them. (hpre at
// *hend is overwritten by the next level's scan.)
// With hbeg_dev the first entry is read from device memory instead; with
// total_copy the total is also written there (the per-level table of
// searches whose level bounds stay on the device).
__global__ void heavy_scan_kernel(const uint32_t* __restrict__ hq, uint32_t hbeg,
                                  const uint32_t* __restrict__ hend,
                                  const QEnt* __restrict__ q, const uint32_t* __restrict__ off,
                                  uint32_t* __restrict__ hpre, uint32_t* __restrict__ total,
                                  const uint32_t* __restrict__ hbeg_dev = nullptr,
                                  uint32_t* __restrict__ total_copy = nullptr) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    if (hbeg_dev) hbeg = *hbeg_dev;
    const uint32_t he = *hend;
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t i0 = hbeg; i0 < he; i0 += blockDim.x) {
        uint32_t i = i0 + threadIdx.x;
        uint32_t c = 0;
        if (i < he) {
            uint32_t v = q[hq[i]].v;
            c = (off[v + 1] - off[v] + kChunkArcs - 1) / kChunkArcs;
        }
        uint32_t x = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wid] = x;
        __syncthreads();
        uint32_t before = carry, tot = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
            if (w < wid) before += wsum[w];
            tot += wsum[w];
        }
        if (i < he) hpre[i] = before + x - c;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        hpre[he] = carry;
        *total = carry;
        if (total_copy) *total_copy = carry;
    }
}

// Heavy entry of chunk task t: the last i in [hbeg, hend) with hpre[i] <= t.
__device__ __forceinline__ uint32_t heavy_of_task(const uint32_t* __restrict__ hpre,
                                                  uint32_t hbeg, uint32_t hend, uint32_t t) {
    uint32_t lo = hbeg, hi = hend;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (hpre[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t wave_or(uint64_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o, 64);
    return x;
}

// Active lanes of slot b when `left` sources remain from the slot-0 lane 0
// of the super-batch on.
__device__ __forceinline__ uint64_t active_lanes(int64_t left, uint32_t b) {
    int64_t r = left - (int64_t)b * kLanes;
    if (r <= 0) return 0;
    return r >= kLanes ? ~0ull : (~0ull >> (kLanes - r));
}

// Items per wavefront task when `count` items are spread over `waves`
// wavefronts: about four tasks each, 1 to `most` items per task. Few items
// per task keep small levels spread over the GPU; many keep the memory
// system busy on large ones.
__host__ __device__ __forceinline__ uint32_t group_size(size_t count, size_t waves,
                                                        uint32_t most = 64) {
    size_t g = (count + 4 * waves - 1) / (4 * waves);
    return (uint32_t)(g < 1 ? 1 : (g > most ? most : g));
}

// Entries per dependency-sweep task at most. Tasks are handed out by
// ticket; on R-MAT-18 groups of 32 beat 16 and 64 by 5-7%, and 4 (too many
// tickets) by 2x.
#ifndef NETBC_SWEEP_GROUP
#define NETBC_SWEEP_GROUP 32
#endif
constexpr uint32_t kSweepGroup = NETBC_SWEEP_GROUP;

// XCD-local work. Workgroups are dispatched to the device's 8 XCDs round-
// robin, so workgroup i runs on XCD i % 8, and each XCD has its own L2.
// Kernels whose work is ordered by slot give XCD k the slots b with
// b % 8 == k, so an XCD's L2 holds the bitmaps of its own slots only.
// Their grids are multiples of 8 (xcd_grid), so every XCD class has waves.
constexpr uint32_t kXcds = 8;
struct XcdWave {
    uint32_t xcd;   // XCD class of this wavefront's workgroup
    size_t w;       // wavefront index within the class
    size_t stride;  // wavefronts in the class
};
__device__ __forceinline__ XcdWave xcd_wave() {
    const size_t wpb = blockDim.x >> 6;
    return XcdWave{blockIdx.x % kXcds, (size_t)(blockIdx.x / kXcds) * wpb + (threadIdx.x >> 6),
                   (size_t)(gridDim.x / kXcds) * wpb};
}
// Slots b < slots with b % 8 == k.
__device__ __forceinline__ uint32_t xcd_slots(uint32_t k, int slots) {
    return (uint32_t)slots > k ? ((uint32_t)slots - k + kXcds - 1) / kXcds : 0;
}

// Levels of at least this many entries are swept in slot order.
constexpr uint32_t kSlotOrderMin = 1u << 16;

// Consecutive 64-pair groups per ticket of the counting bottom-up step on
// large searches (R-MAT 18 all sources: 1.42 s at 1, 1.40 at 2, 1.38 at 4
// and 8, 1.42 at 16; same-address atomics serialise, so fewer tickets wait
// less; on G(2000, 10000), with a few groups per wavefront, 4 per ticket
// took the run from 1.8 to 2.7 ms). A power of two.
#ifndef NETBC_PULL_GROUP
#define NETBC_PULL_GROUP 4
#endif
// Ticket classes of the dependency sweep (a multiple of kXcds).
constexpr uint32_t kTicketClasses = 32;

// Next task of the dependency sweep (wave-uniform): the wavefronts of
// class xc (blockIdx.x % kTicketClasses) take the tasks xc, xc + 32, ...
// from their own counter, so 32 counters share the atomic traffic.
__device__ __forceinline__ uint32_t next_ticket(Ctrl* ctrl, uint32_t xc) {
    uint32_t t = 0;
    if (lane_id() == 0) t = atomicAdd(&ctrl->xticket[xc].ticket, 1u);
    return xc + kTicketClasses * rl32(t, 0);
}
// Called once per wavefront after its last ticket: the last one of its
// class resets the class's counters for the next launch on the stream.
// Every class must hold a workgroup (ticket_grid).
__device__ __forceinline__ void retire_tickets(Ctrl* ctrl, uint32_t xc) {
    const uint32_t blocks =
        gridDim.x / kTicketClasses + (xc < gridDim.x % kTicketClasses ? 1u : 0u);
    const uint32_t waves = blocks * (blockDim.x >> 6);
    if (lane_id() == 0 && atomicAdd(&ctrl->xticket[xc].exits, 1u) == waves - 1) {
        ctrl->xticket[xc].ticket = 0;
        ctrl->xticket[xc].exits = 0;
    }
}

// Work tickets per XCD class k (blockIdx.x % kXcds) for kernels whose work
// is split by XCD (the first kXcds counter pairs of xticket; the sweep's
// retire leaves them at zero): the wavefronts of class k take the class's
// task groups in order from one counter, so the slow groups do not hold up
// a static share. retire_xcd_tickets: once per wavefront after its last
// ticket; the class's last wavefront resets the pair. The grid must be a
// multiple of kXcds (xcd_grid).
__device__ __forceinline__ uint32_t next_xcd_ticket(Ctrl* ctrl, uint32_t k) {
    uint32_t t = 0;
    if (lane_id() == 0) t = atomicAdd(&ctrl->xticket[k].ticket, 1u);
    return rl32(t, 0);
}
__device__ __forceinline__ void retire_xcd_tickets(Ctrl* ctrl, uint32_t k) {
    const uint32_t waves = (gridDim.x / kXcds) * (blockDim.x >> 6);
    if (lane_id() == 0 && atomicAdd(&ctrl->xticket[k].exits, 1u) == waves - 1) {
        ctrl->xticket[k].ticket = 0;
        ctrl->xticket[k].exits = 0;
    }
}

// Exclusive prefix of per-lane counts over the wavefront; total = their sum.
__device__ __forceinline__ uint32_t wave_prefix(uint32_t v, uint32_t& total) {
    const int lane = lane_id();
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    total = rl32(incl, 63);
    return incl - v;
}

// The lane whose items hold position t of the concatenation of the lanes'
// items (pre: exclusive prefix of the lanes' item counts): the last lane c
// with pre[c] <= t. Lanes without items share their prefix with the next
// lane that has some, so the last such lane is the owner; for t past the
// end the result is lane 63.
__device__ __forceinline__ int owner_lane(uint32_t pre, uint32_t t) {
    int lo = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        uint32_t pm = (uint32_t)__shfl((int)pre, lo + step, 64);
        if (pm <= t) lo += step;
    }
    return lo;
}

// Level lookups of the path-count gathers and the dependency sweeps: the
// lanes in which the node of a (slot, node) pair sits at the level an arc
// leads from (gathers) or to (sweeps). Breadth-first searches read one
// bitmap per level (LevelBits). Integer lengths keep the bitmap of every
// distance in a ring of W buffers of `stride` words (W = longest arc + 1,
// slot D mod W for distance D; d0 = D mod W for the current distance D):
// an arc of length len leads from distance D - len (RingBits<true>) or to
// distance D + len (RingBits<false>).
struct LevelBits {
    const uint64_t* __restrict__ bits;
    __device__ __forceinline__ uint64_t lanes(size_t pair, uint64_t) const { return bits[pair]; }
};
template <bool kBack>
struct RingBits {
    const uint64_t* __restrict__ ring;
    size_t stride;
    const uint32_t* __restrict__ len;  // arc lengths, in the order of the arcs walked
    uint32_t d0, W;
    __device__ __forceinline__ uint64_t lanes(size_t pair, uint64_t arc) const {
        const uint32_t l = len[arc];
        uint32_t sl;
        if constexpr (kBack) {
            sl = d0 >= l ? d0 - l : d0 + W - l;
        } else {
            sl = d0 + l;
            if (sl >= W) sl -= W;
        }
        return ring[(size_t)sl * stride + pair];
    }
};

// Path counts of the wavefront's 64 pairs; lane c holds pair c: slot b,
// node x, the lanes U sought and the in-row [i0, i0 + deg) (deg = 0: none).
// The in-rows are walked as one concatenation, 64 arcs at a time; an arc
// u->x hits in the lanes m = prev(u) & U in which u is at level d. Hits are
// then taken in order, so each pair's hits are contiguous and its sums stay
// in registers until its last one: sigma[x][l] = sum of sigma[u][l] over
// its hits with l in m (integers below 2^53: exact in any order, so equal
// to the reference's pushes). With succ, every hit also records that u has
// a successor in lanes m. Returns in lane c the lanes N_c = OR of its hits.
// With kAdd the sums are added atomically (exact: integer partial sums)
// instead of stored, for pairs whose in-row is split over several tasks.
// With kNpr each lane's count of predecessor arcs is also written to npr.
template <bool kAdd, bool kNpr = false, class Prev = LevelBits>
__device__ __forceinline__ uint64_t gather_counts(uint32_t x, uint32_t b, uint64_t U, uint32_t i0,
                                                  uint32_t deg, const uint32_t* __restrict__ in_idx,
                                                  const Prev prev,
                                                  uint64_t* __restrict__ succ,
                                                  double* __restrict__ sig, size_t n,
                                                  unsigned long long& arcs_seen,
                                                  unsigned long long& hits_seen,
                                                  uint32_t* __restrict__ npr = nullptr) {
    const int lane = lane_id();
    uint32_t T;
    const uint32_t pre = wave_prefix(deg, T);
    uint64_t mine = 0;
    int curc = -1;  // pair whose hits are being summed (wave-uniform)
    uint64_t ncur = 0;
    double acc = 0.0;
    uint32_t cnt = 0;  // predecessor arcs in this lane (npr != nullptr)
    auto close_pair = [&]() {
        if (curc < 0) return;
        const uint32_t xs = rl32(x, curc), bs = rl32(b, curc);
        if ((ncur >> lane) & 1) {
            const size_t r = ((size_t)bs * n + xs) * kLanes + lane;
            double* dst = &sig[r];
            if (kAdd) unsafeAtomicAdd(dst, acc);
            else *dst = acc;
            if constexpr (kNpr) {
                if (kAdd) atomicAdd(&npr[r], cnt);
                else npr[r] = cnt;
            }
        }
        if (lane == curc) mine = ncur;
        acc = 0.0;
        cnt = 0;
        ncur = 0;
    };
    if (lane == 0) NETBC_COUNT(arcs_seen, T);
    for (uint64_t t0 = 0; t0 < T; t0 += 64) {
        const uint32_t t = (uint32_t)t0 + lane;
        const int c = owner_lane(pre, t);
        const uint32_t bc = (uint32_t)__shfl((int)b, c, 64);
        const uint32_t ic = (uint32_t)__shfl((int)i0, c, 64);
        const uint32_t pc = (uint32_t)__shfl((int)pre, c, 64);
        const uint64_t Uc = __shfl(U, c, 64);
        uint32_t u = 0;
        uint64_t m = 0;
        if (t < T) {
            const uint32_t arc = ic + (t - pc);
            u = in_idx[arc];
            m = prev.lanes((size_t)bc * n + u, arc) & Uc;
            if (succ && m)
                atomicOr((unsigned long long*)&succ[(size_t)bc * n + u], (unsigned long long)m);
        }
        NETBC_COUNT(hits_seen, m != 0);
        uint64_t hits = __ballot(m != 0);
        while (hits) {
            double val[kGatherAhead];
            int hc[kGatherAhead];
            uint64_t hm[kGatherAhead];
#pragma unroll
            for (int k = 0; k < kGatherAhead; k++) {
                val[k] = 0.0;
                hc[k] = -1;
                hm[k] = 0;
                if (hits) {
                    const int j = __ffsll((unsigned long long)hits) - 1;
                    hits &= hits - 1;
                    const uint32_t uu = rl32(u, j);
                    const uint32_t bb = rl32(bc, j);
                    hm[k] = rl64(m, j);
                    hc[k] = (int)rl32((uint32_t)c, j);
                    if ((hm[k] >> lane) & 1) val[k] = sig[((size_t)bb * n + uu) * kLanes + lane];
                }
            }
#pragma unroll
            for (int k = 0; k < kGatherAhead; k++) {
                if (hc[k] < 0) break;
                if (hc[k] != curc) {
                    close_pair();
                    curc = hc[k];
                }
                ncur |= hm[k];
                acc += val[k];
                if constexpr (kNpr) cnt += (uint32_t)((hm[k] >> lane) & 1);
            }
        }
    }
    close_pair();
    return mine;
}

// Level-queue entries staged per wavefront in LDS (kPullBuf of them, from
// `buf`) and appended a batch at a time: one queue-slot atomic per flush
// instead of one per 64-pair group. Entries of heavy rows also get their
// queue positions listed in hq. cnt is wave-uniform; every lane calls.
struct QueueStage {
    QEnt* buf;
    uint32_t cnt;
    __device__ __forceinline__ void flush(Ctrl* ctrl, QEnt* q, uint32_t qcap, uint32_t* hq,
                                          const uint8_t* heavy_flag) {
        const int lane = lane_id();
        uint32_t at = 0;
        if (lane == 0) at = atomicAdd(&ctrl->qlen, cnt);
        at = rl32(at, 0);
        for (uint32_t i = lane; i < cnt; i += 64) {
            // field by field: a whole-struct copy goes through scratch
            const uint32_t v = buf[i].v, b = buf[i].b;
            const unsigned long long m = buf[i].m;
            if (at + i < qcap) {
                q[at + i].v = v;
                q[at + i].b = b;
                q[at + i].m = m;
                if (heavy_flag[v]) hq[atomicAdd(&ctrl->hlen, 1u)] = at + i;
            } else {
                ctrl->qoverflow = 1;
            }
        }
        cnt = 0;
    }
    // Stages (v, b, m) for the lanes with `has`.
    __device__ __forceinline__ void add(bool has, uint32_t v, uint32_t b, uint64_t m, Ctrl* ctrl,
                                        QEnt* q, uint32_t qcap, uint32_t* hq,
                                        const uint8_t* heavy_flag) {
        const uint64_t bal = __ballot(has);
        if (!bal) return;
        const uint32_t k = (uint32_t)__popcll(bal);
        if (cnt + k > kPullBuf) flush(ctrl, q, qcap, hq, heavy_flag);
        if (has) buf[cnt + __popcll(bal & lanemask_lt(lane_id()))] = QEnt{v, b, (unsigned long long)m};
        cnt += k;
    }
};

// Top-down step of level d: discovery. Every frontier entry (v, b, M)
// offers its lanes to each out-neighbour x not yet reached in them
// (nv = M & ~vis[x]); offers are ORed into cand[x], and x enters the
// touched list on the first offer of the level. The lanes in which v
// offered anything are those in which it has a successor at level d+1
// (succ). Path counts are gathered afterwards (bfs_gather_kernel,
// bfs_hub_kernel). Light rows are taken 64 entries per wavefront and walked
// as one concatenation 64 arcs at a time; heavy rows (listed in hq) as
// tasks of kChunkArcs arcs. Threads also clear the stale level d-1 entries
// of the next-level bitmap. succ may be null (searches without a sweep).
__global__ void bfs_expand_kernel(const QEnt* __restrict__ q, uint32_t qbeg, uint32_t qend,
                                  uint32_t zbeg, uint32_t zend, const uint32_t* __restrict__ hq,
                                  uint32_t hbeg, uint32_t hend, const uint32_t* __restrict__ hpre,
                                  uint32_t hchunks, const uint32_t* __restrict__ off,
                                  const uint32_t* __restrict__ idx,
                                  const uint64_t* __restrict__ vis, uint64_t* __restrict__ cand,
                                  uint64_t* __restrict__ nxt, uint64_t* __restrict__ succ,
                                  uint2* __restrict__ touched, uint32_t* __restrict__ tlen,
                                  const uint8_t* __restrict__ heavy_flag, Ctrl* ctrl,
                                  unsigned long long* __restrict__ front_done, uint32_t gsz,
                                  size_t n) {
    const int lane = lane_id();
    [[maybe_unused]] unsigned long long c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nth = (size_t)gridDim.x * blockDim.x;
    if (tid < (size_t)kSpread) front_done[tid] = 0;  // read by the host before this launch
    for (size_t i = zbeg + tid; i < zend; i += nth) {
        QEnt e = q[i];
        nxt[(size_t)e.b * n + e.v] = 0;
    }
    // Touched nodes are staged per wave in LDS and appended kTouchBuf at a
    // time, so the shared list counter sees one atomic per batch of nodes.
    __shared__ uint2 tbuf[kBlock / 64][kTouchBuf];
    uint2* mybuf = tbuf[threadIdx.x >> 6];
    uint32_t tcnt = 0;  // wave-uniform
    auto flush = [&]() {
        uint32_t at = 0;
        if (lane == 0) at = atomicAdd(tlen, tcnt);
        at = rl32(at, 0);
        for (uint32_t i = lane; i < tcnt; i += 64) touched[at + i] = mybuf[i];
        tcnt = 0;
    };
    const uint64_t lt = lanemask_lt(lane);
    // Offers lanes nv of slot bb to x (ok: this lane holds an arc).
    auto offer = [&](uint32_t x, uint32_t bb, uint64_t nv) {
        bool first = false;
        if (nv) {
            unsigned long long* cp = (unsigned long long*)&cand[(size_t)bb * n + x];
            if ((*cp & nv) != nv) first = atomicOr(cp, (unsigned long long)nv) == 0;
        }
        NETBC_COUNT(c2, nv != 0);
        NETBC_COUNT(c3, __popcll(nv));
        uint64_t bal = __ballot(first);
        if (bal) {
            uint32_t k = (uint32_t)__popcll(bal);
            if (tcnt + k > kTouchBuf) flush();
            if (first) mybuf[tcnt + __popcll(bal & lt)] = make_uint2(x, bb);
            tcnt += k;
        }
    };
    const size_t nq = qend - qbeg;
    const size_t groups = (nq + gsz - 1) / gsz;
    const size_t tasks = groups + hchunks;
    // tasks by ticket (entries' rows differ in length)
    const uint32_t xc = blockIdx.x % kTicketClasses;
    for (size_t w = next_ticket(ctrl, xc); w < tasks; w = next_ticket(ctrl, xc)) {
        if (w < groups) {
            // gsz light entries, their rows as one concatenation.
            const size_t i = qbeg + w * gsz + lane;
            QEnt en{0, 0, 0};
            uint32_t a0 = 0, deg = 0;
            if ((uint32_t)lane < gsz && i < qend) {
                en = q[i];
                if (!heavy_flag[en.v]) {
                    a0 = off[en.v];
                    deg = off[en.v + 1] - a0;
                }
                NETBC_COUNT(c0, 1);
            }
            uint32_t T;
            const uint32_t pre = wave_prefix(deg, T);
            if (lane == 0) NETBC_COUNT(c1, T);
            for (uint64_t t0 = 0; t0 < T; t0 += 64) {
                const uint32_t t = (uint32_t)t0 + lane;
                const int c = owner_lane(pre, t);
                const uint32_t bc = (uint32_t)__shfl((int)en.b, c, 64);
                const uint32_t vc = (uint32_t)__shfl((int)en.v, c, 64);
                const uint32_t ac = (uint32_t)__shfl((int)a0, c, 64);
                const uint32_t pc = (uint32_t)__shfl((int)pre, c, 64);
                const uint64_t mc = __shfl(en.m, c, 64);
                uint32_t x = 0;
                uint64_t nv = 0;
                if (t < T) {
                    x = idx[ac + (t - pc)];
                    nv = mc & ~vis[(size_t)bc * n + x];
                }
                offer(x, bc, nv);
                // successor lanes of each entry: OR over each run of lanes
                // with the same owner (owners ascend with the lane)
                uint64_t run = nv;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    uint64_t y = __shfl_up(run, o, 64);
                    int cy = __shfl_up(c, o, 64);
                    if (lane >= o && cy == c) run |= y;
                }
                const int cn = __shfl_down(c, 1, 64);
                if (succ && (lane == 63 || cn != c) && run)
                    atomicOr((unsigned long long*)&succ[(size_t)bc * n + vc], (unsigned long long)run);
            }
        } else {
            // One chunk of a heavy row.
            const uint32_t t = (uint32_t)(w - groups);
            const uint32_t h = heavy_of_task(hpre, hbeg, hend, t);
            const QEnt en = q[hq[h]];
            const uint64_t a0 = (uint64_t)off[en.v] + (uint64_t)(t - hpre[h]) * kChunkArcs;
            const uint64_t a1 = min(a0 + kChunkArcs, (uint64_t)off[en.v + 1]);
            const size_t base = (size_t)en.b * n;
            uint64_t offered = 0;
            if (lane == 0) NETBC_COUNT(c1, a1 - a0);
            for (uint64_t c = a0; c < a1; c += 64) {
                const uint64_t a = c + lane;
                uint32_t x = 0;
                uint64_t nv = 0;
                if (a < a1) {
                    x = idx[a];
                    nv = en.m & ~vis[base + x];
                }
                offered |= nv;
                offer(x, en.b, nv);
            }
            offered = wave_or(offered);
            if (succ && lane == 0 && offered)
                atomicOr((unsigned long long*)&succ[base + en.v], (unsigned long long)offered);
        }
    }
    retire_tickets(ctrl, xc);
    if (tcnt) flush();
    NETBC_FLUSH(ctrl, 0, c0);
    NETBC_FLUSH(ctrl, 1, c1);
    NETBC_FLUSH(ctrl, 2, c2);
    NETBC_FLUSH(ctrl, 3, c3);
}

// Hub tasks of level d: task (slot b, chunk of hub x's in-arcs) gathers,
// over the chunk's arcs u->x, sigma[u][l] for the lanes l of N held by u at
// level d, and adds the partial sum into sigma[x][l] (integers below 2^53:
// the partial sums are exact and their order does not matter). After a
// top-down expansion N = cand[x], the lanes in which x was reached. In a
// bottom-up step (bottom_up != 0) N = F_b & ~vis[x] (F_b: the lanes still
// searching, flanes): the lanes found are
// ORed into cand[x], x enters the touched list with its first lanes, and
// the in-neighbours record their successor lanes (succ). A wavefront takes
// gsz tasks of one slot and walks the arcs of those with N != 0 as one
// concatenation (gather_counts); slots are spread over the XCDs (xcd_wave)
// and each XCD class takes its task groups in order by ticket.
// kNpr: predecessor-arc counts as well (npr).
template <bool kNpr>
__global__ void bfs_hub_kernel(const uint4* __restrict__ hub_task, uint32_t chunks, int slots,
                               const uint32_t* __restrict__ in_idx,
                               const uint64_t* __restrict__ cur, uint64_t* __restrict__ cand,
                               double* __restrict__ sig, int bottom_up,
                               const uint64_t* __restrict__ vis, uint64_t* __restrict__ succ,
                               uint2* __restrict__ touched, uint32_t* __restrict__ tlen,
                               const uint64_t* __restrict__ flanes, Ctrl* ctrl, uint32_t gsz,
                               size_t n, uint32_t* __restrict__ npr) {
    const int lane = lane_id();
    const uint64_t lt = lanemask_lt(lane);
    [[maybe_unused]] unsigned long long c4 = 0, c5 = 0, c6 = 0;
    const XcdWave xw = xcd_wave();
    const size_t gps = (chunks + gsz - 1) / gsz;  // task groups per slot
    const size_t groups = (size_t)xcd_slots(xw.xcd, slots) * gps;
    for (size_t g = next_xcd_ticket(ctrl, xw.xcd); g < groups;
         g = next_xcd_ticket(ctrl, xw.xcd)) {
        const uint32_t bs = xw.xcd + kXcds * (uint32_t)(g / gps);
        const uint32_t c = (uint32_t)(g % gps) * gsz + lane;
        uint32_t b = 0;
        uint4 ht = make_uint4(0, 0, 0, 0);
        uint64_t N = 0;
        if ((uint32_t)lane < gsz && c < chunks) {
            b = bs;
            ht = hub_task[c];
            const size_t p = (size_t)b * n + ht.x;
            N = bottom_up ? (flanes[b] & ~vis[p]) : cand[p];
        }
        const uint32_t deg = N ? ht.z - ht.y : 0;
        NETBC_COUNT(c6, N != 0);
        const uint64_t got = gather_counts<true, kNpr>(ht.x, b, N, ht.y, deg, in_idx, LevelBits{cur},
                                                 bottom_up ? succ : nullptr, sig, n, c4, c5,
                                                 npr);
        if (bottom_up) {
            bool first = false;
            if (got)
                first = atomicOr((unsigned long long*)&cand[(size_t)b * n + ht.x],
                                 (unsigned long long)got) == 0;
            const uint64_t bal = __ballot(first);
            if (bal) {
                uint32_t pos = 0;
                if (lane == 0) pos = atomicAdd(tlen, (uint32_t)__popcll(bal));
                pos = rl32(pos, 0) + (uint32_t)__popcll(bal & lt);
                if (first) touched[pos] = make_uint2(ht.x, b);
            }
        }
    }
    retire_xcd_tickets(ctrl, xw.xcd);
    NETBC_FLUSH(ctrl, 12, c4);
    NETBC_FLUSH(ctrl, 13, c5);
    NETBC_FLUSH(ctrl, 14, c6);
}

// Bottom-up hub tasks of a search for reachability only (no path counts):
// a wavefront per task (slot b, chunk of hub x's in-arcs) ORs, 64 arcs at a
// time, the lanes of N = F_b & ~vis[x] held at level d by the in-neighbours,
// and stops once every lane of N is found. The lanes found are ORed into
// cand[x] and x enters the touched list with its first lanes.
__global__ void reach_hub_kernel(const uint4* __restrict__ hub_task, uint32_t chunks, int slots,
                                 const uint32_t* __restrict__ in_idx,
                                 const uint64_t* __restrict__ cur, uint64_t* __restrict__ cand,
                                 const uint64_t* __restrict__ vis, uint2* __restrict__ touched,
                                 uint32_t* __restrict__ tlen, const uint64_t* __restrict__ flanes,
                                 size_t n) {
    const int lane = lane_id();
    // touched pairs staged per wavefront, appended 64 at a time
    __shared__ uint2 tbuf[kBlock / 64][64];
    uint2* mybuf = tbuf[threadIdx.x >> 6];
    uint32_t cnt = 0;  // wave-uniform
    auto flush = [&]() {
        uint32_t at = 0;
        if (lane == 0) at = atomicAdd(tlen, cnt);
        at = rl32(at, 0);
        if ((uint32_t)lane < cnt) touched[at + lane] = mybuf[lane];
        cnt = 0;
    };
    const size_t tasks = (size_t)chunks * slots;
    for (size_t w = global_wave(); w < tasks; w += wave_count()) {
        const uint32_t b = (uint32_t)(w / chunks);
        const uint4 ht = hub_task[w % chunks];
        const size_t p = (size_t)b * n + ht.x;
        const uint64_t N = flanes[b] & ~vis[p];
        if (!N) continue;
        const uint64_t* crow = cur + (size_t)b * n;
        uint64_t got = 0;
        for (uint64_t a = ht.y; a < ht.z && got != N; a += 64) {
            uint64_t m = 0;
            if (a + lane < ht.z) m = crow[in_idx[a + lane]] & N;
            got |= wave_or(m);
        }
        bool first = false;
        if (lane == 0 && got)
            first = atomicOr((unsigned long long*)&cand[p], (unsigned long long)got) == 0;
        if (__ballot(first)) {
            if (cnt == 64) flush();
            if (lane == 0) mybuf[cnt] = make_uint2(ht.x, b);
            cnt++;
        }
    }
    if (cnt) flush();
}

// Lanes of each slot that hold entries of a level (the lanes whose search
// is still running): fl[b] |= the masks of the level's entries in slot b,
// ORed per block in LDS and added with one atomic per block and slot.
__global__ void frontier_lanes_kernel(const QEnt* __restrict__ q, uint32_t qbeg, uint32_t qend,
                                      uint64_t* __restrict__ fl, int slots) {
    extern __shared__ unsigned long long fl_lds[];
    for (int b = threadIdx.x; b < slots; b += blockDim.x) fl_lds[b] = 0;
    __syncthreads();
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    for (size_t i = qbeg + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < qend; i += nth) {
        const QEnt e = q[i];
        atomicOr(&fl_lds[e.b], (unsigned long long)e.m);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < slots; b += blockDim.x)
        if (fl_lds[b]) atomicOr((unsigned long long*)&fl[b], fl_lds[b]);
}

// Closes level d: every touched pair (x, b) takes its new lanes N = cand[x]
// (the OR of all offers), which become visited, form the level d+1 bitmap
// and join the level queue as (x, b, N). The path counts of non-hub pairs
// are gathered here from their in-arcs (gather_counts with U = N; after a
// bottom-up step only hubs are listed, whose counts come from
// bfs_hub_kernel). Also counts the out-arcs of the new entries and the
// in-arcs of pairs whose active lanes are now all reached (the direction
// choice of the next step). Without kCount (reachability only) no path
// counts are gathered; kNpr: predecessor-arc counts as well (npr).
template <bool kCount, bool kNpr = false>
__global__ void bfs_gather_kernel(const uint2* __restrict__ touched,
                                  const uint32_t* __restrict__ tlen_cur,
                                  uint32_t* __restrict__ tlen_next,
                                  const uint32_t* __restrict__ in_off,
                                  const uint32_t* __restrict__ in_idx,
                                  const uint32_t* __restrict__ out_off,
                                  const uint64_t* __restrict__ cur, uint64_t* __restrict__ cand,
                                  uint64_t* __restrict__ vis, uint64_t* __restrict__ nxt,
                                  double* __restrict__ sig, const uint8_t* __restrict__ hub_flag,
                                  const uint8_t* __restrict__ heavy_flag, QEnt* __restrict__ q,
                                  uint32_t qcap, uint32_t* __restrict__ hq, Ctrl* ctrl,
                                  int64_t left, unsigned long long* __restrict__ front_next,
                                  size_t n, uint32_t* __restrict__ npr) {
    const int lane = lane_id();
    if (blockIdx.x == 0 && threadIdx.x == 0) *tlen_next = 0;
    const uint32_t T = *tlen_cur;
    unsigned long long fa = 0, closed = 0;
    [[maybe_unused]] unsigned long long c4 = 0, c5 = 0;
    const uint32_t gsz = group_size(T, wave_count());
    // new entries staged in dynamic LDS (kPullLds bytes per block)
    extern __shared__ unsigned long long pull_lds[];
    QueueStage stage{reinterpret_cast<QEnt*>(pull_lds) + (threadIdx.x >> 6) * kPullBuf, 0};
    // groups of touched pairs by ticket when counting (their in-rows differ
    // in length); reachability-only groups are cheap, and a static share
    // costs less than a ticket atomic per group
    const uint32_t xc = blockIdx.x % kTicketClasses;
    for (size_t k = kCount ? next_ticket(ctrl, xc) : global_wave(); k * gsz < T;
         k = kCount ? next_ticket(ctrl, xc) : k + wave_count()) {
        const size_t t0 = k * gsz;
        const size_t t = t0 + lane;
        const bool ok = (uint32_t)lane < gsz && t < T;
        const uint2 te = ok ? touched[t] : make_uint2(0, 0);
        const size_t p = (size_t)te.y * n + te.x;
        uint64_t N = 0;
        uint32_t i0 = 0, i1 = 0, deg = 0;
        if (ok) {
            N = cand[p];
            i0 = in_off[te.x];
            i1 = in_off[te.x + 1];
            if (!hub_flag[te.x]) deg = i1 - i0;
        }
        if constexpr (kCount)
            (void)gather_counts<false, kNpr>(te.x, te.y, N, i0, deg, in_idx, LevelBits{cur}, nullptr, sig,
                                             n, c4, c5, npr);
        if (ok) {
            cand[p] = 0;
            const uint64_t vn = vis[p] | N;
            vis[p] = vn;
            nxt[p] = N;
            fa += out_off[te.x + 1] - out_off[te.x];
            if (vn == active_lanes(left, te.y)) closed += i1 - i0;
        }
        stage.add(ok, te.x, te.y, N, ctrl, q, qcap, hq, heavy_flag);
    }
    if (kCount) retire_tickets(ctrl, xc);
    if (stage.cnt) stage.flush(ctrl, q, qcap, hq, heavy_flag);
    NETBC_FLUSH(ctrl, 10, c4);
    NETBC_FLUSH(ctrl, 11, c5);
    fa = wave_sum_u64(fa);
    closed = wave_sum_u64(closed);
    if (lane == 0) {
        if (fa) atomicAdd(&front_next[blockIdx.x % kSpread], fa);
        if (closed) atomicAdd(&ctrl->closed_arcs[blockIdx.x % kSpread], closed);
    }
}

// Bottom-up step of level d (direction-optimising search). Every (slot b,
// node x) pair that is not a hub, has in-arcs and has lanes not yet reached
// among those still searching, U = F_b & ~vis[x] (F_b: the lanes with
// level-d entries, flanes), gathers over its in-arcs the lanes held by
// level-d in-neighbours and their path counts (gather_counts, recording
// successor lanes in succ); the pairs reached join level d+1 here. A
// wavefront takes 64 consecutive pairs of one slot; slots are spread over
// the XCDs (xcd_wave). Hubs are left to bfs_hub_kernel and
// bfs_gather_kernel. Each XCD class takes its 64-pair groups in order by
// ticket (groups differ in cost: their in-rows and hits). Threads also
// clear the level d-1 entries of the next bitmap (AND-NOT, commuting with
// the OR of the new lanes: a node's lanes at levels d-1 and d+1 are
// disjoint). Without kCount (reachability only)
// each pair walks its own in-row and stops at the first in-neighbours that
// cover every lane sought. kNpr: predecessor-arc counts as well (npr).
template <bool kCount, bool kNpr = false>
__global__ void bfs_pull_kernel(const QEnt* qz, uint32_t zbeg, uint32_t zend,
                                const uint32_t* __restrict__ in_off,
                                const uint32_t* __restrict__ in_idx,
                                const uint32_t* __restrict__ out_off, uint64_t* __restrict__ vis,
                                const uint64_t* __restrict__ cur, uint64_t* __restrict__ nxt,
                                uint64_t* __restrict__ succ, double* __restrict__ sig,
                                const uint8_t* __restrict__ hub_flag,
                                const uint8_t* __restrict__ heavy_flag, QEnt* q, uint32_t qcap,
                                uint32_t* __restrict__ hq, Ctrl* ctrl,
                                unsigned long long* __restrict__ front_next,
                                unsigned long long* __restrict__ front_done,
                                const uint64_t* __restrict__ flanes, int slots, int64_t left,
                                size_t n, uint32_t* __restrict__ npr) {
    const int lane = lane_id();
    size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nth = (size_t)gridDim.x * blockDim.x;
    if (tid < (size_t)kSpread) front_done[tid] = 0;  // read by the host before this launch
    for (size_t i = zbeg + tid; i < zend; i += nth) {
        QEnt e = qz[i];
        atomicAnd((unsigned long long*)&nxt[(size_t)e.b * n + e.v], ~e.m);
    }
    unsigned long long fa = 0, closed = 0, bytes = 0;
    [[maybe_unused]] unsigned long long c4 = 0, c5 = 0, c6 = 0;
    // new entries staged in dynamic LDS (kPullLds bytes per block)
    extern __shared__ unsigned long long pull_lds[];
    QueueStage stage{reinterpret_cast<QEnt*>(pull_lds) + (threadIdx.x >> 6) * kPullBuf, 0};
    const XcdWave xw = xcd_wave();
    const size_t gps = (n + 63) / 64;  // 64-pair groups per slot
    const size_t groups = (size_t)xcd_slots(xw.xcd, slots) * gps;
    // the counting step takes its groups by ticket (they differ in cost);
    // reachability-only groups are cheap, and a static share costs less
    // than a ticket atomic per group
    // counting: NETBC_PULL_GROUP consecutive groups per ticket when there
    // are at least 8 groups per wavefront (small searches keep one, so that
    // their few groups spread over the wavefronts)
    const size_t G = groups >= xw.stride * 8 ? NETBC_PULL_GROUP : 1;
    for (size_t g = kCount ? (size_t)next_xcd_ticket(ctrl, xw.xcd) * G : xw.w; g < groups;
         g = !kCount                ? g + xw.stride
             : ((g + 1) & (G - 1)) ? g + 1
                                   : (size_t)next_xcd_ticket(ctrl, xw.xcd) * G) {
        const uint32_t b = xw.xcd + kXcds * (uint32_t)(g / gps);
        const uint64_t F = flanes[b];  // lanes still searching in this slot
        if (!F) continue;
        const uint32_t x0 = (uint32_t)(g % gps) * 64;
        const bool ok = x0 + (uint32_t)lane < n;
        const uint32_t x = ok ? x0 + lane : 0;
        const size_t p = (size_t)b * n + x;
        uint64_t U = 0, V = 0;
        uint32_t i0 = 0, deg = 0, odeg = 0;
        if (ok) {
            // independent loads, issued together
            V = vis[p];
            const uint8_t hub = hub_flag[x];
            const uint32_t a0 = in_off[x], a1 = in_off[x + 1];
            odeg = out_off[x + 1] - out_off[x];
            U = F & ~V;
            if (U && !hub) {
                i0 = a0;
                deg = a1 - a0;
            }
        }
        NETBC_COUNT(c6, deg != 0);
        uint64_t mine = 0;
        if constexpr (kCount) {
            mine = gather_counts<false, kNpr>(x, b, U, i0, deg, in_idx, LevelBits{cur}, succ, sig, n, c4, c5,
                                              npr);
        } else if (deg) {
            const uint64_t* crow = cur + (size_t)b * n;
            const uint32_t* row = in_idx + i0;
            uint32_t a = 0;
            for (; a + 4 <= deg && mine != U; a += 4)
                mine |= (crow[row[a]] | crow[row[a + 1]] | crow[row[a + 2]] | crow[row[a + 3]]) & U;
            for (; a < deg && mine != U; a++) mine |= crow[row[a]] & U;
            bytes += 8 + 12ull * a;
        }
        if (!kCount && ok) bytes += 8 + (mine ? 40 : 0);
        if (mine) {
            const uint64_t vn = V | mine;
            vis[p] = vn;
            atomicOr((unsigned long long*)&nxt[p], (unsigned long long)mine);
            fa += odeg;
            if (vn == active_lanes(left, b)) closed += deg;
        }
        stage.add(mine != 0, x, b, mine, ctrl, q, qcap, hq, heavy_flag);
    }
    if (stage.cnt) stage.flush(ctrl, q, qcap, hq, heavy_flag);
    if (kCount) retire_xcd_tickets(ctrl, xw.xcd);
    NETBC_FLUSH(ctrl, 8, c4);
    NETBC_FLUSH(ctrl, 9, c5);
    NETBC_FLUSH(ctrl, 15, c6);
    fa = wave_sum_u64(fa);
    closed = wave_sum_u64(closed);
    if constexpr (!kCount) bytes = wave_sum_u64(bytes);
    if (lane == 0) {
        if (fa) atomicAdd(&front_next[blockIdx.x % kSpread], fa);
        if (closed) atomicAdd(&ctrl->closed_arcs[blockIdx.x % kSpread], closed);
        if (!kCount && bytes) atomicAdd(&ctrl->reach_bytes[blockIdx.x % kSpread], bytes);
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
    // (t + delta) / sigma, t = 1 when the node is a target (every node is,
    // except in runs for sources and targets, path_centrality.rs:457-459)
    __device__ __forceinline__ static double coef(double t, double delta, double sigma) {
        return (t + delta) / sigma;
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
    __device__ __forceinline__ static float coef(double t, double delta, double sigma) {
        return (float)((t + delta) / sigma);
    }
};

// Cross-lane moves for edge_flush. v_permlane32_swap_b32 exchanges the
// upper 32 lanes of its first operand with the lower 32 lanes of its
// second, v_permlane16_swap_b32 the upper 16 with the lower 16 lanes of
// each half-wave, so after the swap the two registers hold, in lane L, the
// operand pair's values of lanes L and L + d (L below d within each 2d
// group: first operand) or L - d and L (second operand). swap_add returns
// their sum. DPP quad_perm and row_ror move values within rows of 16 lanes
// (row_ror:k: lane L reads lane L - k mod 16).
template <int d>
__device__ __forceinline__ double swap_add(double a, double b) {
    static_assert(d == 16 || d == 32, "half-wave or quarter-wave swap");
    const unsigned alo = (unsigned)__double2loint(a), ahi = (unsigned)__double2hiint(a);
    const unsigned blo = (unsigned)__double2loint(b), bhi = (unsigned)__double2hiint(b);
    if constexpr (d == 32) {
        const auto lo = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
        const auto hi = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
        return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
    } else {
        const auto lo = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
        const auto hi = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
        return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
    }
}

constexpr int kDppXor1 = 0xB1;   // quad_perm [1,0,3,2]: lane L reads L ^ 1
constexpr int kDppXor2 = 0x4E;   // quad_perm [2,3,0,1]: lane L reads L ^ 2
constexpr int kDppRor8 = 0x128;  // row_ror:8: lane L reads L ^ 8
constexpr int kDppRor12 = 0x12C; // row_ror:12: lane L reads L + 4 (mod 16)

template <int kCtrl>
__device__ __forceinline__ double dpp_f64(double x) {
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), kCtrl, 0xF, 0xF, false);
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), kCtrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Edge credit of up to K arcs taken by one wave (reach.rs:287-309: arc
// v -> w on a shortest path carries sigma_v / sigma_w * (1 + delta_w) per
// source). c[h] holds lane l's credit of hit h (0 where the lane's source
// does not use the arc), ak[h] the arc (kNone: no hit) and bk[h] its slot.
// The 64 lanes of every hit are added by a fixed reduce-scatter: log2(K)
// halving exchanges (lane bits 5, 4, 3) leave lane L with hit
// L >> (6 - log2 K) summed over the lanes that differ from L in those bits,
// and sums over the low bits complete it in the lanes whose low bits are
// zero, which write it. The exchanges use register swaps and DPP moves,
// not LDS. The totals are added to eacc[slot * arcs + arc]: one sweep
// launch meets a (slot, arc) pair at most once, so the update needs no
// atomic.
template <int K>
__device__ __forceinline__ void edge_flush(double (&c)[K], const uint32_t (&ak)[K],
                                           const uint32_t (&bk)[K], double* __restrict__ eacc,
                                           size_t arcs) {
    static_assert(K == 1 || K == 2 || K == 4 || K == 8, "hits per flush: 1, 2, 4 or 8");
    constexpr int kLog = K == 1 ? 0 : K == 2 ? 1 : K == 4 ? 2 : 3;
    const int lane = lane_id();
    if constexpr (kLog >= 1) {
#pragma unroll
        for (int i = 0; i < K / 2; i++) c[i] = swap_add<32>(c[i], c[i + K / 2]);
    }
    if constexpr (kLog >= 2) {
#pragma unroll
        for (int i = 0; i < K / 4; i++) c[i] = swap_add<16>(c[i], c[i + K / 4]);
    }
    if constexpr (kLog >= 3) {
        const bool hi = (lane & 8) != 0;
        const double keep = hi ? c[1] : c[0];
        const double send = hi ? c[0] : c[1];
        c[0] = keep + dpp_f64<kDppRor8>(send);
    }
    double x = c[0];
    x += dpp_f64<kDppXor1>(x);                       // lane bit 0
    x += dpp_f64<kDppXor2>(x);                       // bit 1
    x += dpp_f64<kDppRor12>(x);                      // bit 2 (exact where bit 2 is 0)
    if constexpr (kLog <= 2) x += dpp_f64<kDppRor8>(x);  // bit 3
    if constexpr (kLog <= 1) x = swap_add<16>(x, x);    // bit 4
    if constexpr (kLog == 0) x = swap_add<32>(x, x);    // bit 5
    c[0] = x;
    const int h = lane >> (6 - kLog);
    uint32_t a = kNone, b = 0;
#pragma unroll
    for (int i = 0; i < K; i++)
        if (h == i) {
            a = ak[i];
            b = bk[i];
        }
    if ((lane & ((64 >> kLog) - 1)) == 0 && a != kNone) eacc[(size_t)b * arcs + a] += c[0];
}

// Adds the coefficients of the successors of `en` (next: their lanes, see
// bfs_backward_kernel) among arcs [a0, a1) to dv, per lane in CSR order; up to kAhead
// rows are loaded before they are added. With kEdge, each arc's credit
// sv * coef (sv: the lane's path count of en) goes to eacc.
template <class C, bool kEdge, class Next>
__device__ __forceinline__ void sweep_arcs(const QEnt& en, uint64_t a0, uint64_t a1,
                                           const uint32_t* __restrict__ idx, const Next next,
                                           const C* __restrict__ coef, size_t n, Dep<C>& dv,
                                           double sv, double* __restrict__ eacc, size_t arcs,
                                           unsigned long long& c6, unsigned long long& c7) {
    const int lane = lane_id();
    const size_t base = (size_t)en.b * n;
    for (uint64_t c = a0; c < a1; c += 64) {
        const uint64_t a = c + lane;
        bool ok = a < a1;
        uint32_t x = ok ? idx[a] : 0;
        uint64_t mm = ok ? (next.lanes(base + x, a) & en.m) : 0;
        NETBC_COUNT(c6, mm != 0);
        NETBC_COUNT(c7, __popcll(mm));
        uint64_t bal = __ballot(mm != 0);
        while (bal) {
            C val[kAhead];
            bool use[kAhead];
            uint32_t ak[kAhead], bk[kAhead];
#pragma unroll
            for (int u = 0; u < kAhead; u++) {
                use[u] = false;
                val[u] = 0;
                ak[u] = kNone;
                bk[u] = en.b;
                if (bal) {
                    int j = __ffsll((unsigned long long)bal) - 1;
                    bal &= bal - 1;
                    uint32_t xx = rl32(x, j);
                    uint64_t m2 = rl64(mm, j);
                    ak[u] = (uint32_t)(c + (uint32_t)j);
                    use[u] = (m2 >> lane) & 1;
                    if (use[u]) val[u] = coef[(base + xx) * kLanes + lane];
                }
            }
#pragma unroll
            for (int u = 0; u < kAhead; u++)
                if (use[u]) dv.add(val[u]);
            if constexpr (kEdge) {
                double cv[kAhead];
#pragma unroll
                for (int u = 0; u < kAhead; u++) cv[u] = use[u] ? sv * (double)val[u] : 0.0;
                edge_flush<kAhead>(cv, ak, bk, eacc, arcs);
            }
        }
    }
}

// Adds the sum over the wavefront's lanes of x (one fixed butterfly) to *dst.
// Every lane calls. Free-order breadth-first runs add each entry's
// dependencies into one word per (slot, node) this way; a node has one
// entry per level, and levels are swept one launch at a time, so the adds
// to a word come in level order and their sum is deterministic.
__device__ __forceinline__ void add_lane_sum(double* dst, double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane_id() == 0 && x != 0.0) unsafeAtomicAdd(dst, x);
}

// Wavefronts per SIMD the sweep's registers are capped for: 8 (64 VGPRs,
// up to 16 bytes per lane spilled) for node scores, 587 -> 566 ms per
// R-MAT-18 run; edge runs keep their 65 (7 waves: capped, they spill 9-14
// VGPRs and run 2% slower).
constexpr int sweep_waves(bool edge) { return edge ? 1 : 8; }

// Dependencies of the level-d entries; `nx` gives the lanes in which an
// arc's head is a successor (breadth-first: the level d+1 bitmap; integer
// lengths: the bitmap of distance D + len in the distance ring). With
// rebuild, W is turned from the level d+2 bitmap into level d's (lanes are
// disjoint, so the OR and AND-NOT updates commute); levels kept from the
// forward search need no rebuild. A wavefront takes up to 64 light entries:
// the rows of those with a successor in their lanes (succ) are walked as
// one concatenation, 64 arcs at a time, and the successors' coefficients
// are added per lane in CSR order (hits are taken in order, so one row's
// sum is complete before the next row's starts); the other entries have no
// dependency (delta = 0). Heavy rows (listed in hq) are split into chunk
// tasks that each write a per-lane partial sum (part), added in chunk order
// by bfs_combine_kernel; with split == 0 every row is summed whole. With
// kEdge every arc's credit is also added to eacc (edge_flush). With kQuirk
// (NetworkX's subset edge split, path_centrality.rs:465-474) the rows hold
// E_w instead of coefficients, E_w = (1 + de_w) / sigma_w for a target and
// de_w / np_w otherwise (np_w: w's predecessor arcs, npr), the arc v -> w
// carries sigma_v E_w or E_w, and de_v is their sum; rows are whole. With
// bsum (free-order runs) each finished entry also adds the sum of its
// lanes' dependencies to bsum[b][v] (add_lane_sum; bfs_combine_kernel does
// the same for heavy rows).
template <class C, bool kEdge, bool kQuirk = false, class Next = LevelBits>
__global__ void __launch_bounds__(1024, sweep_waves(kEdge)) bfs_backward_kernel(const QEnt* __restrict__ q, uint32_t qbeg, uint32_t qend,
                                    uint32_t abeg, uint32_t aend, int last_level, int rebuild,
                                    const uint32_t* __restrict__ hq, uint32_t hbeg, uint32_t hend,
                                    const uint32_t* __restrict__ hpre, uint32_t hchunks, int split,
                                    const uint8_t* __restrict__ heavy_flag,
                                    double* __restrict__ part,
                                    const uint32_t* __restrict__ off,
                                    const uint32_t* __restrict__ idx, const Next nx,
                                    uint64_t* __restrict__ W, const uint64_t* __restrict__ succ,
                                    double* __restrict__ sig, C* __restrict__ coef, Ctrl* ctrl,
                                    uint32_t gsz, size_t n, double* __restrict__ eacc,
                                    size_t arcs, const uint8_t* __restrict__ target,
                                    uint32_t* __restrict__ npr, double* __restrict__ bsum,
                                    const uint32_t* __restrict__ perm) {
    const int lane = lane_id();
    double smax = 0.0;
    [[maybe_unused]] unsigned long long c4 = 0, c5 = 0, c6 = 0, c7 = 0;
    if (rebuild) {
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
    split = split && !last_level && !kQuirk;
    // delta, coefficient and path-count bookkeeping of lane l of entry (v, b, M);
    // returns the lane's dependency (0 outside M)
    auto finish = [&](uint32_t v, uint32_t b, uint64_t M, double s, double dvv,
                      [[maybe_unused]] double dqv) -> double {
        double dep = 0.0;
        if ((M >> lane) & 1) {
            const size_t r = ((size_t)b * n + v) * kLanes + lane;
            if constexpr (kQuirk) {
                const double de = s * dvv + dqv;
                const uint32_t np = npr[r];
                npr[r] = 0;  // rows start at zero for the next super-batch
                coef[r] = (C)(target_of(target, v) != 0.0 ? (1.0 + de) / s
                                                          : (np ? de / (double)np : 0.0));
                sig[r] = de;
                dep = de;
            } else {
                const double delta = s * dvv;
                coef[r] = Dep<C>::coef(target_of(target, v), delta, s);
                // with bsum the dependency goes into the (slot, node) sum
                // alone; nothing reads the row afterwards
                if (!bsum) sig[r] = delta;
                dep = delta;
            }
            smax = fmax(smax, s);
        }
        return dep;
    };
    const size_t nq = qend - qbeg;
    const size_t groups = (nq + gsz - 1) / gsz;
    const size_t tasks = groups + (split ? hchunks : 0);
    // Tasks are handed out by ticket, the next one fetched while the
    // current one runs.
    const uint32_t xc = blockIdx.x % kTicketClasses;
    uint32_t next = next_ticket(ctrl, xc);
    while (next < tasks) {
        const size_t w = next;
        next = next_ticket(ctrl, xc);
        if (w >= groups) {
            // One chunk of a heavy row: its partial sum.
            uint32_t t = (uint32_t)(w - groups);
            uint32_t h = heavy_of_task(hpre, hbeg, hend, t);
            QEnt en = q[hq[h]];
            const uint64_t a0 = (uint64_t)off[en.v] + (uint64_t)(t - hpre[h]) * kChunkArcs;
            const uint64_t a1 = min(a0 + kChunkArcs, (uint64_t)off[en.v + 1]);
            Dep<C> dv;
            if (succ[(size_t)en.b * n + en.v] & en.m) {
                // the row's path counts stay in place until bfs_combine_kernel
                double sv = 0.0;
                if (kEdge && ((en.m >> lane) & 1))
                    sv = sig[((size_t)en.b * n + en.v) * kLanes + lane];
                sweep_arcs<C, kEdge>(en, a0, a1, idx, nx, coef, n, dv, sv, eacc, arcs, c6, c7);
            }
            part[t * kLanes + lane] = dv.value();
            continue;
        }
        const size_t i = qbeg + w * gsz + lane;
        QEnt en{0, 0, 0};
        bool live = false;  // this lane's entry is finished by this wave
        uint32_t a0 = 0, deg = 0;
        if ((uint32_t)lane < gsz && i < qend) {
            en = q[perm ? perm[i] : i];
            live = !(split && heavy_flag[en.v]);
            // rows without a successor in any of their lanes add nothing
            if (live && !last_level && (succ[(size_t)en.b * n + en.v] & en.m)) {
                a0 = off[en.v];
                deg = off[en.v + 1] - a0;
            }
        }
        NETBC_COUNT(c4, live);
        uint32_t T;
        const uint32_t pre = wave_prefix(deg, T);
        if (lane == 0) NETBC_COUNT(c5, T);
        bool done = !live;
        int cur = -1;       // entry whose successors are being summed (wave-uniform)
        Dep<C> dv, dq;      // dq: kQuirk, successors that are not targets
        double sv = 0.0;    // its path counts, loaded when it starts
        auto close_entry = [&]() {
            if (cur < 0) return;
            const uint32_t vv = rl32(en.v, cur), bb = rl32(en.b, cur);
            const double dep = finish(vv, bb, rl64(en.m, cur), sv, dv.value(), dq.value());
            if (bsum) add_lane_sum(&bsum[(size_t)bb * n + vv], dep);
            if (lane == cur) done = true;
            dv = Dep<C>();
            dq = Dep<C>();
        };
        auto open_entry = [&](int c) {
            cur = c;
            const uint32_t vv = rl32(en.v, c), bb = rl32(en.b, c);
            const uint64_t mm = rl64(en.m, c);
            sv = ((mm >> lane) & 1) ? sig[((size_t)bb * n + vv) * kLanes + lane] : 0.0;
        };
        for (uint64_t t0 = 0; t0 < T; t0 += 64) {
            const uint32_t t = (uint32_t)t0 + lane;
            const int c = owner_lane(pre, t);
            const uint32_t bc = (uint32_t)__shfl((int)en.b, c, 64);
            const uint32_t ac = (uint32_t)__shfl((int)a0, c, 64);
            const uint32_t pc = (uint32_t)__shfl((int)pre, c, 64);
            const uint64_t mc = __shfl(en.m, c, 64);
            uint32_t x = 0, arc = 0;
            uint64_t mm = 0;
            if (t < T) {
                arc = ac + (t - pc);
                x = idx[arc];
                mm = nx.lanes((size_t)bc * n + x, arc) & mc;
            }
            NETBC_COUNT(c6, mm != 0);
            NETBC_COUNT(c7, __popcll(mm));
            uint64_t bal = __ballot(mm != 0);
            while (bal) {
                C val[kAhead];
                bool use[kAhead];
                int hc[kAhead];
                uint32_t ak[kAhead], bk[kAhead];
                [[maybe_unused]] bool tk[kAhead];
#pragma unroll
                for (int k = 0; k < kAhead; k++) {
                    use[k] = false;
                    val[k] = 0;
                    hc[k] = -1;
                    ak[k] = kNone;
                    bk[k] = 0;
                    tk[k] = true;
                    if (bal) {
                        const int j = __ffsll((unsigned long long)bal) - 1;
                        bal &= bal - 1;
                        const uint32_t xx = rl32(x, j), bb = rl32(bc, j);
                        const uint64_t m2 = rl64(mm, j);
                        hc[k] = (int)rl32((uint32_t)c, j);
                        ak[k] = rl32(arc, j);
                        bk[k] = bb;
                        if constexpr (kQuirk) tk[k] = target_of(target, xx) != 0.0;
                        use[k] = (m2 >> lane) & 1;
                        if (use[k]) val[k] = coef[((size_t)bb * n + xx) * kLanes + lane];
                    }
                }
                [[maybe_unused]] double cv[kAhead];
#pragma unroll
                for (int k = 0; k < kAhead; k++) {
                    cv[k] = 0.0;
                    if (hc[k] < 0) continue;
                    if (hc[k] != cur) {
                        close_entry();
                        open_entry(hc[k]);
                    }
                    if (use[k]) {
                        if (!kQuirk || tk[k]) {
                            dv.add(val[k]);
                            if constexpr (kEdge) cv[k] = sv * (double)val[k];
                        } else {
                            dq.add(val[k]);
                            cv[k] = (double)val[k];
                        }
                    }
                }
                if constexpr (kEdge) edge_flush<kAhead>(cv, ak, bk, eacc, arcs);
            }
        }
        close_entry();
        // The rest have no dependency; their path counts are loaded kAhead
        // entries at a time.
        uint64_t rest = __ballot(!done);
        while (rest) {
            double s[kAhead];
            int ec[kAhead];
#pragma unroll
            for (int k = 0; k < kAhead; k++) {
                s[k] = 0.0;
                ec[k] = -1;
                if (rest) {
                    const int j = __ffsll((unsigned long long)rest) - 1;
                    rest &= rest - 1;
                    ec[k] = j;
                    const uint32_t vv = rl32(en.v, j), bb = rl32(en.b, j);
                    const uint64_t mm = rl64(en.m, j);
                    if ((mm >> lane) & 1) s[k] = sig[((size_t)bb * n + vv) * kLanes + lane];
                }
            }
#pragma unroll
            for (int k = 0; k < kAhead; k++) {
                if (ec[k] < 0) break;
                finish(rl32(en.v, ec[k]), rl32(en.b, ec[k]), rl64(en.m, ec[k]), s[k],
                       Dep<C>().value(), 0.0);
            }
        }
    }
    retire_tickets(ctrl, xc);
    note_sigma(ctrl, smax);
    NETBC_FLUSH(ctrl, 4, c4);
    NETBC_FLUSH(ctrl, 5, c5);
    NETBC_FLUSH(ctrl, 6, c6);
    NETBC_FLUSH(ctrl, 7, c7);
}

// Slot order of a level's queue entries for the dependency sweep: perm
// lists the positions [qb, qe) of the level's entries grouped by slot (in
// no particular order within a slot; the sweep's results do not depend on
// the order of a level's entries). The sweep hands out its tasks in
// increasing order, so the wavefronts at work at any time share one or two
// slots and their successor lookups (random reads of the next level's
// words, 2 MB per slot at n = 262,144) stay in each XCD's L2.
// Counts per slot (cnt zeroed before).
__global__ void slot_count_kernel(const QEnt* __restrict__ q, uint32_t qb, uint32_t qe, int slots,
                                  uint32_t* __restrict__ cnt) {
    extern __shared__ uint32_t hist[];
    for (int b = threadIdx.x; b < slots; b += blockDim.x) hist[b] = 0;
    __syncthreads();
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = qb + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < qe; i += stride)
        atomicAdd(&hist[q[i].b], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < slots; b += blockDim.x)
        if (hist[b]) atomicAdd(&cnt[b], hist[b]);
}

// cnt[b] = the first position of slot b (exclusive prefix over slots; one
// workgroup of 1024 threads, slots <= 1024).
__global__ void __launch_bounds__(1024) slot_scan_kernel(uint32_t* __restrict__ cnt, int slots) {
    __shared__ uint32_t wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t v = t < slots ? cnt[t] : 0;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (int k = 0; k < w; k++) before += wsum[k];
    if (t < slots) cnt[t] = before + incl - v;
}

// perm[qb + cursor[b]++] = i for every entry i of [qb, qe) in slot b; each
// workgroup reserves one range per slot per round of blockDim entries.
__global__ void slot_scatter_kernel(const QEnt* __restrict__ q, uint32_t qb, uint32_t qe,
                                    int slots, uint32_t* __restrict__ cursor,
                                    uint32_t* __restrict__ perm) {
    extern __shared__ uint32_t lds[];
    uint32_t* hist = lds;          // entries per slot this round
    uint32_t* base = lds + slots;  // their first position
    const size_t per_round = (size_t)gridDim.x * blockDim.x;
    for (size_t r0 = qb + (size_t)blockIdx.x * blockDim.x; r0 < qe; r0 += per_round) {
        for (int b = threadIdx.x; b < slots; b += blockDim.x) hist[b] = 0;
        __syncthreads();
        const size_t i = r0 + threadIdx.x;
        uint32_t b = 0, rank = 0;
        if (i < qe) {
            b = q[i].b;
            rank = atomicAdd(&hist[b], 1u);
        }
        __syncthreads();
        for (int s = threadIdx.x; s < slots; s += blockDim.x)
            if (hist[s]) base[s] = atomicAdd(&cursor[s], hist[s]);
        __syncthreads();
        if (i < qe) perm[qb + base[b] + rank] = (uint32_t)i;
        __syncthreads();
    }
}

// Finishes the heavy rows of a level: dv = the chunk partial sums added in
// chunk order, then the row's dependencies and coefficients.
template <class C>
__global__ void bfs_combine_kernel(const QEnt* __restrict__ q, const uint32_t* __restrict__ hq,
                                   uint32_t hbeg, uint32_t hend, const uint32_t* __restrict__ hpre,
                                   uint32_t hchunks, const double* __restrict__ part,
                                   const uint32_t* __restrict__ off, double* __restrict__ sig,
                                   C* __restrict__ coef, Ctrl* ctrl, size_t n,
                                   const uint8_t* __restrict__ target,
                                   double* __restrict__ bsum) {
    const int lane = lane_id();
    double smax = 0.0;
    for (size_t h = global_wave(); h < hend - hbeg; h += wave_count()) {
        QEnt en = q[hq[hbeg + h]];
        // The next level's prefix starts at hpre[hend], so the last row of
        // this level ends at the level's chunk total.
        const uint32_t c0 = hpre[hbeg + h];
        const uint32_t c1 = hbeg + h + 1 < hend ? hpre[hbeg + h + 1] : hchunks;
        double dv = 0.0;
        for (uint32_t c = c0; c < c1; c++) dv += part[(size_t)c * kLanes + lane];
        double dep = 0.0;
        if ((en.m >> lane) & 1) {
            size_t r = ((size_t)en.b * n + en.v) * kLanes + lane;
            double s = sig[r];
            double delta = s * dv;
            coef[r] = Dep<C>::coef(target_of(target, en.v), delta, s);
            if (!bsum) sig[r] = delta;
            dep = delta;
            smax = fmax(smax, s);
        }
        if (bsum) add_lane_sum(&bsum[(size_t)en.b * n + en.v], dep);
    }
    note_sigma(ctrl, smax);
}

// ---- bucketed searches: integer lengths (Dial) and real lengths -------------

// Distances and lengths of the bucketed searches. Integer lengths: a bucket
// per distance (Dial, betweenness.rs:214-295), on the path iff equal. Real
// lengths: buckets of width delta <= half the shortest arc, so an arc always
// leads to a later bucket and every tentative distance in the lowest
// pending bucket is final; a distance is the least offer it received, and
// an arc is on a shortest path iff its offer ties that distance within a
// relative 1e-10 (approx_eq, betweenness.rs:117-120, 297-368).
struct IntPath {
    using Dist = uint32_t;
    using Len = uint32_t;
    using Sum = uint64_t;  // distance + length without wrapping
    [[maybe_unused]] static constexpr uint64_t kInfBits = kNone;
    __device__ static Sum add(Dist d, Len l) { return (Sum)d + l; }
    __device__ static bool reached(Dist d) { return d != kNone; }
    // buckets of `width` distances (the tuned bucket width; 1: a bucket per distance)
    __device__ static uint64_t bucket(Sum x, double width) {
        return width == 1.0 ? x : x / (uint64_t)width;
    }
    __device__ static bool tie(Sum via, Dist d) { return via == (Sum)d; }
    __device__ static void offer(Dist* p, Sum v) { atomicMin(p, (uint32_t)v); }
};

struct RealPath {
    using Dist = double;
    using Len = double;
    using Sum = double;
    static constexpr uint64_t kInfBits = 0x7FF0000000000000ull;  // +infinity
    __device__ static Sum add(Dist d, Len l) { return d + l; }
    __device__ static bool reached(Dist d) { return d < INFINITY; }
    __device__ static uint64_t bucket(Sum x, double delta) { return (uint64_t)(x / delta); }
    __device__ static bool tie(Sum via, Dist d) {
        const double m = fabs(via) > fabs(d) ? fabs(via) : fabs(d);
        return fabs(via - d) <= 1e-10 * m;
    }
    // Non-negative doubles order as their bit patterns.
    __device__ static void offer(Dist* p, Sum v) {
        atomicMin((unsigned long long*)p, (unsigned long long)__double_as_longlong(v));
    }
};

// Real lengths with ties only between equal lengths (Tie::Exact,
// paths/mod.rs:96-104, NetworkX's rule).
struct RealPathExact : RealPath {
    __device__ static bool tie(Sum via, Dist d) { return via == d; }
};

// ---- integer lengths (Dial / delta-stepping with unit buckets) -------------

// Integer lengths, forward step D (Dial's algorithm with exact integer
// distances, betweenness.rs:214-295, run for 64 sources per slot). Each
// (slot, node) keeps a bitmask of the buckets (distance mod W, W = max
// length + 1) in which some lane holds a tentative distance; all tentative
// distances lie in [D, D + W), so bucket D mod W holds exactly distance D.
//
// dial_scan_kernel lists the (slot, node) pairs with bucket D mod W set,
// with the node's out-row bounds (the step starts on the row while the
// pair's distance row is in flight), and clears it, records whether any bucket is still set anywhere, and lowers
// next_bucket to the first bucket after D still set at some pair (the
// pending buckets of a pair lie in (D, D + W), so a residue r names bucket
// D + ((r - D) mod W)); the step adds the buckets it offers to. Pairs of
// rows longer than kLongRow are listed from the list's end instead. It runs
// one workgroup per CU (with a workgroup per resident slot the wavefronts'
// appends to the list's one counter made it 3x slower), each lane loading
// kScanUnroll bucket words at once. The wavefronts advance through the
// pairs together and append as they go, so the list, and the level queue
// built from it, stay near slot-major order, which the step and the sweep
// rely on for L2 hits (one append per workgroup at its end measured 2%
// slower overall for that reason).
constexpr int kScanUnroll = 4;
// Pairs whose node has more out-arcs than this are listed from the end of
// the list (count in tlong), and the step takes them first: one wavefront
// walks a row, so a long row taken last would end the launch alone (BA
// n = 200,000, 32,768 sources: 1.844 s with no such rows, 1.80 s from 512
// arcs; from 128 or 256 the sweep, fed in the step's order, slowed more
// than the step gained).
#ifndef NETBC_LONG_ROW
#define NETBC_LONG_ROW 512
#endif
constexpr uint32_t kLongRow = NETBC_LONG_ROW;
__global__ void dial_scan_kernel(uint64_t* __restrict__ pend, int words, int W, uint32_t D,
                                 size_t count, const uint32_t* __restrict__ off,
                                 uint4* __restrict__ list,
                                 uint32_t* __restrict__ llen, uint32_t* __restrict__ llen_next,
                                 uint32_t* __restrict__ tlong, uint32_t* __restrict__ tlong_next,
                                 uint32_t* __restrict__ left, uint32_t* __restrict__ left_next,
                                 unsigned long long* __restrict__ next_bucket,
                                 unsigned long long* __restrict__ next_bucket_next, size_t n) {
    const int kb = (int)(D % (uint32_t)W);
    __shared__ uint4 sbuf[kBlock / 64][kTouchBuf];
    const int lane = lane_id();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *llen_next = 0;
        *tlong_next = 0;
        *left_next = 0;
        *next_bucket_next = ~0ull;
    }
    uint4* mybuf = sbuf[threadIdx.x >> 6];
    uint32_t cnt = 0;
    auto flush = [&]() {
        uint32_t at = 0;
        if (lane == 0) at = atomicAdd(llen, cnt);
        at = rl32(at, 0);
        for (uint32_t i = lane; i < cnt; i += 64) list[at + i] = mybuf[i];
        cnt = 0;
    };
    const uint64_t lt = lanemask_lt(lane);
    const unsigned long long bit = 1ull << (kb & 63);
    const int kw = kb >> 6;
    // residues above kb within its word, and below it
    const uint64_t above = (kb & 63) == 63 ? 0ull : (~0ull << ((kb & 63) + 1));
    const uint64_t below = bit - 1;
    int any = 0;
    uint32_t ahead = kNone;  // least (r - kb) mod W over pending residues r
    for (size_t w0 = global_wave() * 64 * kScanUnroll; w0 < count;
         w0 += wave_count() * 64 * kScanUnroll) {
        // the bucket words of kScanUnroll pairs per lane loaded together
        uint64_t wk[kScanUnroll];
#pragma unroll
        for (int u = 0; u < kScanUnroll; u++) {
            const size_t i = w0 + (size_t)u * 64 + lane;
            wk[u] = i < count ? pend[i * words + kw] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < kScanUnroll; u++) {
            const size_t i = w0 + (size_t)u * 64 + lane;
            bool hit = false;
            if (i < count) {
                uint64_t* p = &pend[i * words];
                uint64_t word = wk[u];
                hit = (word & bit) != 0;
                if (hit) p[kw] = word & ~bit;
                word &= ~bit;
                uint32_t hi = kNone, lo = kNone;  // first residue above kb / at all below it
                if (word & above) hi = (uint32_t)(kw * 64 + __ffsll((unsigned long long)(word & above)) - 1);
                if (word & below) lo = (uint32_t)(kw * 64 + __ffsll((unsigned long long)(word & below)) - 1);
                for (int j = 0; j < words; j++) {
                    if (j == kw) continue;
                    const uint64_t wj = p[j];
                    if (!wj) continue;
                    const uint32_t r = (uint32_t)(j * 64 + __ffsll((unsigned long long)wj) - 1);
                    if (j > kw) hi = min(hi, r);
                    else lo = min(lo, r);
                }
                if (hi != kNone) ahead = min(ahead, hi - (uint32_t)kb);
                else if (lo != kNone) ahead = min(ahead, lo + (uint32_t)W - (uint32_t)kb);
                any |= (hi != kNone || lo != kNone);
            }
            uint64_t bal = __ballot(hit);
            if (bal) {
                uint32_t x = 0, o0 = 0, o1 = 0;
                if (hit) {
                    x = (uint32_t)(i % n);
                    o0 = off[x];
                    o1 = off[x + 1];
                }
                const bool lng = hit && o1 - o0 > kLongRow;
                const uint64_t lb = __ballot(lng);
                if (lb) {
                    uint32_t at = 0;
                    if (lane == 0) at = atomicAdd(tlong, (uint32_t)__popcll(lb));
                    at = rl32(at, 0);
                    if (lng)
                        list[count - 1 - (at + (uint32_t)__popcll(lb & lt))] =
                            make_uint4(x, (uint32_t)(i / n), o0, o1);
                    bal &= ~lb;
                }
                const uint32_t k = (uint32_t)__popcll(bal);
                if (k) {
                    if (cnt + k > kTouchBuf) flush();
                    if (hit && !lng)
                        mybuf[cnt + __popcll(bal & lt)] =
                            make_uint4(x, (uint32_t)(i / n), o0, o1);
                    cnt += k;
                }
            }
        }
    }
    if (cnt) flush();
    for (int o = 32; o > 0; o >>= 1) ahead = min(ahead, (uint32_t)__shfl_xor((int)ahead, o, 64));
    // plain reads first: most wavefronts change neither value (a stale read
    // only lets an atomic through)
    if (lane == 0 && ahead != kNone &&
        (unsigned long long)D + ahead < *(volatile unsigned long long*)next_bucket)
        atomicMin(next_bucket, (unsigned long long)D + ahead);
    if (__syncthreads_or(any) && threadIdx.x == 0 && !*(volatile uint32_t*)left)
        atomicOr(left, 1u);
}

// The Dial sweep packs the (head, lane) pairs of entries of at most
// kPackLanes lanes (dial_sweep_packed; wider entries measured faster
// unpacked), the undirected step those of every entry (dial_step_packed:
// BA n = 200,000, 16,384 sources 0.935 s with 16 lanes, 0.922 s with 32,
// 0.88 s with all; its unpacked walk loads one neighbour, about 9 of 64
// lanes, per load).
#ifndef NETBC_PACK_LANES
#define NETBC_PACK_LANES 16
#endif
constexpr uint32_t kPackLanes = NETBC_PACK_LANES;
// 64-pair loads per packed window.
#ifndef NETBC_PACK_ROUNDS
#define NETBC_PACK_ROUNDS 2
#endif
constexpr int kPackRounds = NETBC_PACK_ROUNDS;

// The k-th lane of M in lane k < popcount(M) (0 elsewhere).
__device__ __forceinline__ uint32_t select_lanes(uint64_t M) {
    const int lane = lane_id();
    uint32_t sel = 0;
    if ((uint64_t)lane < (uint64_t)__popcll(M)) {
        uint64_t m = M;
        for (int k = 0; k < lane; k++) m &= m - 1;
        sel = (uint32_t)__ffsll((unsigned long long)m) - 1;
    }
    return sel;
}

// The undirected Dial step's row walk for an entry of lanes M with its
// (neighbour, lane) pairs packed 64 to a load: each
// load covers 64 / |M| whole neighbours, kPackRounds loads a window. A
// pair whose lane is already settled at the neighbour is a predecessor
// candidate (its path count is loaded when its distance ties, with
// `settle`), any other takes the offer dist[x] + length by atomic minimum
// and marks the offer's bucket (one atomic per distinct neighbour and
// bucket). Each lane of M then adds its predecessors' path counts in CSR
// order, as the unpacked walk does, reading them from the window's values
// staged in the wavefront's LDS row `stage` (64 * kPackRounds doubles; one
// LDS read per value instead of two lane shuffles). With `preloaded` the first 64 arcs
// (y0, ln0, vw0) are loaded by the caller. kWalk 0: the undirected row
// (both); 1: an in-row (predecessors only); 2: an out-row (offers only).
// Every lane calls.
template <class P, bool kCount, int kWalk = 0>
__device__ __forceinline__ void dial_step_packed(
    uint64_t a0, uint64_t a1, bool preloaded, uint32_t y0, typename P::Len ln0, uint64_t vw0,
    uint64_t M,
    size_t base, typename P::Dist dx, int W, int words, double delta, bool settle,
    const uint32_t* __restrict__ idx, const typename P::Len* __restrict__ len_arr,
    const uint64_t* __restrict__ vis, typename P::Dist* __restrict__ dist,
    const double* __restrict__ sig, uint64_t* __restrict__ pend, double& acc, uint32_t& npred,
    unsigned long long& least, double* __restrict__ stage) {
    using Dist = typename P::Dist;
    using Len = typename P::Len;
    using Sum = typename P::Sum;
    const int lane = lane_id();
    const uint32_t mc = (uint32_t)__popcll(M);
    const uint32_t sel = select_lanes(M);
    const bool mine = (M >> lane) & 1;
    const uint32_t rank = (uint32_t)__popcll(M & lanemask_lt(lane));
    const uint32_t A = 64 / mc;  // whole neighbours per 64-pair load
    const uint32_t grp = (uint32_t)lane / mc, gi = (uint32_t)lane - grp * mc;
    const uint32_t ls = (uint32_t)__shfl((int)sel, (int)gi, 64);  // this pair's lane
    const Dist d0 = __shfl(dx, (int)ls, 64);                      // its distance at x
    for (uint64_t c = a0; c < a1; c += 64) {
        const uint64_t a = c + lane;
        const bool ok = a < a1;
        uint32_t y = y0;
        Len ln = ln0;
        uint64_t vw = vw0;
        if (c != a0 || !preloaded) {
            y = ok ? idx[a] : 0;
            ln = ok ? len_arr[a] : Len(0);
            vw = ok ? vis[base + y] : 0;
        }
        const uint32_t na = (uint32_t)min<uint64_t>(64, a1 - c);
        for (uint32_t w0 = 0; w0 < na; w0 += A * kPackRounds) {
            Dist du[kPackRounds];
            bool pr[kPackRounds], hd[kPackRounds];
            size_t row[kPackRounds];
            Len lp[kPackRounds];
            uint32_t yr[kPackRounds];
#pragma unroll
            for (int r = 0; r < kPackRounds; r++) {
                const uint32_t j = w0 + (uint32_t)r * A + grp;  // neighbour in this chunk
                const bool valid = grp < A && j < na;
                const int jj = (int)min(j, 63u);
                yr[r] = (uint32_t)__shfl((int)y, jj, 64);
                lp[r] = __shfl(ln, jj, 64);
                const uint64_t vj = __shfl(vw, jj, 64);
                const bool settled = (vj >> ls) & 1;
                pr[r] = kWalk != 2 && valid && kCount && settle && settled;
                hd[r] = kWalk != 1 && valid && !settled;
                row[r] = (base + yr[r]) * kLanes + ls;
                du[r] = pr[r] || hd[r] ? dist[row[r]] : Dist(0);
            }
            double su[kPackRounds];
            uint64_t tm[kPackRounds];
#pragma unroll
            for (int r = 0; r < kPackRounds; r++) {
                const bool tie = pr[r] && P::tie(P::add(du[r], lp[r]), d0);
                su[r] = tie ? sig[row[r]] : 0.0;
                tm[r] = __ballot(tie);
                bool offered = false;
                uint32_t kk = 0;
                const Sum nd = P::add(d0, lp[r]);
                if (hd[r] && nd < (Sum)du[r]) {
                    P::offer(&dist[row[r]], nd);
                    offered = true;
                    const uint64_t bnd = P::bucket(nd, delta);
                    kk = (uint32_t)(bnd % (uint64_t)W);
                    least = min(least, (unsigned long long)bnd);
                }
                // mark the offered buckets: one atomic per distinct
                // (neighbour, bucket) of this load
                uint64_t pending = __ballot(offered);
                while (pending) {
                    const int f = __ffsll((unsigned long long)pending) - 1;
                    const uint32_t kf = rl32(kk, f);
                    const uint32_t gf = (uint32_t)f / mc;
                    pending &= ~__ballot(offered && kk == kf && grp == gf);
                    if (lane == f)
                        atomicOr((unsigned long long*)&pend[(base + yr[r]) * words + (kf >> 6)],
                                 1ull << (kf & 63));
                }
            }
            // each lane of M adds its predecessors' path counts in CSR order
            uint64_t any_tie = 0;
#pragma unroll
            for (int r = 0; r < kPackRounds; r++) any_tie |= tm[r];
            if (any_tie) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int r = 0; r < kPackRounds; r++) stage[r * 64 + lane] = su[r];
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                for (int r = 0; r < kPackRounds; r++) {
                    for (uint32_t g = 0; g < A; g++) {
                        if (w0 + (uint32_t)r * A + g >= na) break;
                        const int src = (int)(g * mc + rank) & 63;
                        if (mine && ((tm[r] >> src) & 1)) {
                            acc += stage[r * 64 + src];
                            npred++;
                        }
                    }
                }
            }
        }
    }
}

// Settles and relaxes the listed pairs of bucket D (P: IntPath or
// RealPath). The lanes of x whose tentative distance lies in bucket D are
// final (M). Their path counts are pulled over in-arcs u->x of length len
// from settled lanes whose offer dist[u] + len ties dist[x], and (x, b, M)
// joins the level queue (both skipped for the sources at D = 0). Then every
// out-arc x->y offers dist[x] + len to the lanes of y not yet settled, by
// atomic minimum, and marks the offer's bucket mod W at y. Offers land in
// later buckets, so settled distances never change here. Without kCount
// (distances only) no path counts are pulled. With kSym (undirected: the
// in-row of x is its out-row) the row is walked once, its (neighbour, lane)
// pairs packed 64 to a load (dial_step_packed): per neighbour y, the lanes
// of M already settled at y are the predecessor candidates and the others
// take offers, so one visited word and one distance row of y serve both
// (lanes at y settled in this same bucket are neither: every length is
// longer than a bucket). Directed graphs walk the in-row (predecessors),
// then the out-row (offers), with the pairs packed the same way.
template <class P, bool kCount = true, bool kSym = false>
__global__ void dial_step_kernel(const uint4* __restrict__ list, const uint32_t* __restrict__ llen,
                                 const uint32_t* __restrict__ tlong, size_t cap, uint32_t D, int W, int words, int first_step, double delta,
                                 uint64_t* __restrict__ pend,
                                 const uint32_t* __restrict__ in_off,
                                 const uint32_t* __restrict__ in_idx,
                                 const typename P::Len* __restrict__ in_len,
                                 const uint32_t* __restrict__ off,
                                 const uint32_t* __restrict__ idx,
                                 const typename P::Len* __restrict__ len_arr,
                                 typename P::Dist* __restrict__ dist, uint64_t* __restrict__ vis,
                                 double* __restrict__ sig, QEnt* __restrict__ q, uint32_t qcap,
                                 Ctrl* ctrl, unsigned long long* __restrict__ next_bucket,
                                 size_t n, uint32_t* __restrict__ npr) {
    using Dist = typename P::Dist;
    using Len = typename P::Len;
    __shared__ QEnt qbuf[kBlock / 64][kTouchBuf / 2];
    // per wavefront: the packed path counts of a window (dial_step_packed)
    __shared__ double stage_rows[kBlock / 64][64 * kPackRounds];
    double* stage = stage_rows[threadIdx.x >> 6];
    unsigned long long least = ~0ull;  // least bucket offered by this lane
    const int lane = lane_id();
    QEnt* mybuf = qbuf[threadIdx.x >> 6];
    uint32_t cnt = 0;
    auto flush = [&]() {
        uint32_t at = 0;
        if (lane == 0) at = atomicAdd(&ctrl->qlen, cnt);
        at = rl32(at, 0);
        for (uint32_t i = lane; i < cnt; i += 64) {
            if (at + i < qcap) q[at + i] = mybuf[i];
            else ctrl->qoverflow = 1;
        }
        cnt = 0;
    };
    double smax = 0.0;
    [[maybe_unused]] unsigned long long c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    // long rows (listed from the end) first
    const uint32_t TL = *tlong;
    const uint32_t T = *llen + TL;
    // entries by ticket when counting (their rows differ widely in length);
    // distances-only entries are cheap, and a static share costs less than
    // a ticket atomic per entry
    const uint32_t xc = blockIdx.x % kTicketClasses;
    // counting: kDialGroup consecutive entries per ticket when there are at
    // least 8 entries per wavefront
    const size_t G = T >= wave_count() * 8 ? kDialGroup : 1;
    for (size_t t = kCount ? (size_t)next_ticket(ctrl, xc) * G : global_wave(); t < T;
         t = !kCount                ? t + wave_count()
             : ((t + 1) & (G - 1)) ? t + 1
                                   : (size_t)next_ticket(ctrl, xc) * G) {
        const uint4 te = t < TL ? list[cap - 1 - t] : list[t - TL];  // node, slot, out-row bounds
        const uint32_t x = te.x;
        const size_t base = (size_t)te.y * n;
        // undirected: the row's first 64 arcs and their visited words are
        // loaded while the distance row is in flight
        uint32_t y0 = 0;
        Len ln0 = Len(0);
        uint64_t vw0 = 0;
        const bool ok0 = (uint64_t)te.z + lane < te.w;
        if constexpr (kSym) {
            if (ok0) {
                y0 = idx[te.z + lane];
                ln0 = len_arr[te.z + lane];
            }
        }
        const Dist dx = dist[(base + x) * kLanes + lane];
        if constexpr (kSym) {
            if (ok0) vw0 = vis[base + y0];
        }
        const uint64_t M = __ballot(P::reached(dx) && P::bucket(dx, delta) == D);
        if (M == 0) continue;
        if (lane == 0) {
            NETBC_COUNT(c0, 1);
            NETBC_COUNT(c1, in_off[x + 1] - in_off[x]);
            NETBC_COUNT(c2, off[x + 1] - off[x]);
        }
        const bool mine = (M >> lane) & 1;
        if constexpr (kSym) {
            double acc = 0.0;
            uint32_t npred = 0;  // predecessor arcs in this lane
            const uint64_t a0 = te.z, a1 = te.w;
            dial_step_packed<P, kCount>(a0, a1, true, y0, ln0, vw0, M, base, dx, W, words, delta,
                                        !first_step, idx, len_arr, vis, dist, sig, pend, acc,
                                        npred, least, stage);
            if (!first_step) {
                if (kCount && mine) {
                    sig[(base + x) * kLanes + lane] = acc;
                    if (npr) npr[(base + x) * kLanes + lane] = npred;
                    smax = fmax(smax, acc);
                }
                if (lane == 0) vis[base + x] |= M;
                if (cnt + 1 > kTouchBuf / 2) flush();
                if (lane == 0) mybuf[cnt] = QEnt{x, te.y, (unsigned long long)M};
                cnt += 1;
                // a long row's entry goes to the queue at once (the sweep
                // takes it first too)
                if (te.w - te.z > kLongRow) flush();
            }
            continue;
        }
        if (!first_step) {
            double acc = 0.0;
            uint32_t npred = 0;  // predecessor arcs in this lane
            // predecessors over the in-row (pairs packed as for undirected rows)
            if (kCount)
                dial_step_packed<P, kCount, 1>(in_off[x], in_off[x + 1], false, 0, Len(0), 0, M,
                                               base, dx, W, words, delta, true, in_idx, in_len,
                                               vis, dist, sig, pend, acc, npred, least, stage);
            if (kCount && mine) {
                sig[(base + x) * kLanes + lane] = acc;
                if (npr) npr[(base + x) * kLanes + lane] = npred;
                smax = fmax(smax, acc);
            }
            if (lane == 0) vis[base + x] |= M;
            if (cnt + 1 > kTouchBuf / 2) flush();
            if (lane == 0) mybuf[cnt] = QEnt{x, te.y, (unsigned long long)M};
            cnt += 1;
            if (te.w - te.z > kLongRow) flush();
        }
        // offers over the out-row
        double acc_unused = 0.0;
        uint32_t npred_unused = 0;
        dial_step_packed<P, kCount, 2>(te.z, te.w, false, 0, Len(0), 0, M, base, dx, W, words,
                                       delta, false, idx, len_arr, vis, dist, sig, pend, acc_unused,
                                       npred_unused, least, stage);
    }
    if (kCount) retire_tickets(ctrl, xc);
    if (cnt) flush();
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long other = (unsigned long long)__shfl_xor((long long)least, o, 64);
        least = min(least, other);
    }
    if (lane == 0 && least != ~0ull && least < *(volatile unsigned long long*)next_bucket)
        atomicMin(next_bucket, least);
    note_sigma(ctrl, smax);
    NETBC_FLUSH(ctrl, 0, c0);
    NETBC_FLUSH(ctrl, 1, c1);
    NETBC_FLUSH(ctrl, 2, c2);
    NETBC_FLUSH(ctrl, 3, c3);
}

// The Dial sweep of one entry (lanes M, base of its slot, own distances
// dv0 in its lanes) whose row is [a0, a1), with the (head, lane) pairs of
// up to 128 / |M| heads at a time packed into two wave-wide loads: pair p
// holds head p / |M| and the (p mod |M|)-th lane of M. All the pairs'
// distances are loaded, then the coefficients of the pairs on a shortest
// path; each lane of M then adds its tying heads' coefficients in CSR
// order, as the unpacked walk does (one load per head there, with 9 of 64
// lanes in use on BA n = 200,000), reading them from the window's
// coefficients staged in the wavefront's LDS row `stage` (64 * kPackRounds
// values). Every lane calls.
template <class C, class P>
__device__ __forceinline__ void dial_sweep_packed(uint64_t a0, uint64_t a1, uint64_t M,
                                                  size_t base, typename P::Dist dv0,
                                                  const uint32_t* __restrict__ idx,
                                                  const typename P::Len* __restrict__ len_arr,
                                                  const typename P::Dist* __restrict__ dist,
                                                  const uint64_t* __restrict__ vis,
                                                  const C* __restrict__ coef, Dep<C>& dv,
                                                  C* __restrict__ stage) {
    using Dist = typename P::Dist;
    using Len = typename P::Len;
    const int lane = lane_id();
    const uint32_t mc = (uint32_t)__popcll(M);
    const uint32_t sel = select_lanes(M);
    const bool mine = (M >> lane) & 1;
    const uint32_t rank = (uint32_t)__popcll(M & lanemask_lt(lane));
    const uint32_t A = 64 * kPackRounds / mc;  // heads per packed window
    for (uint64_t c = a0; c < a1; c += 64) {
        const uint64_t a = c + lane;
        const bool ok = a < a1;
        const uint32_t x = ok ? idx[a] : 0;
        const Len len = ok ? len_arr[a] : Len(0);
        const uint64_t mm = ok ? (vis[base + x] & M) : 0;
        const uint32_t na = (uint32_t)min<uint64_t>(64, a1 - c);
        for (uint32_t w0 = 0; w0 < na; w0 += A) {
            const uint32_t wa = min(A, na - w0);
            const uint32_t pairs = wa * mc;
            Dist dxp[kPackRounds];
            C cv[kPackRounds];
            bool use[kPackRounds];
            size_t row[kPackRounds];
            Len lp[kPackRounds];
            uint32_t lsrc[kPackRounds];
#pragma unroll
            for (int r = 0; r < kPackRounds; r++) {
                const uint32_t p = (uint32_t)r * 64 + (uint32_t)lane;
                const uint32_t ai = p / mc;
                const uint32_t ri = p - ai * mc;
                const uint32_t head = min(w0 + ai, 63u);
                lsrc[r] = (uint32_t)__shfl((int)sel, (int)min(ri, 63u), 64);
                const uint32_t xa = (uint32_t)__shfl((int)x, (int)head, 64);
                lp[r] = __shfl(len, (int)head, 64);
                const uint64_t ma = __shfl(mm, (int)head, 64);
                use[r] = p < pairs && ((ma >> lsrc[r]) & 1);
                row[r] = (base + xa) * kLanes + lsrc[r];
                dxp[r] = use[r] ? dist[row[r]] : Dist(0);
            }
            uint64_t tm[kPackRounds];
#pragma unroll
            for (int r = 0; r < kPackRounds; r++) {
                const Dist d0 = __shfl(dv0, (int)lsrc[r], 64);
                const bool tie = use[r] && P::tie(P::add(d0, lp[r]), dxp[r]);
                cv[r] = tie ? coef[row[r]] : C(0);
                tm[r] = __ballot(tie);
            }
            // each lane of M adds its tying heads in CSR order
            // the packed coefficients staged in this wavefront's LDS row
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int r = 0; r < kPackRounds; r++) stage[r * 64 + lane] = cv[r];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (uint32_t j = 0; j < wa; j++) {
                const uint32_t qp = j * mc + rank;
                uint64_t t = tm[0];
#pragma unroll
                for (int r = 1; r < kPackRounds; r++)
                    if ((qp >> 6) == (uint32_t)r) t = tm[r];
                if (mine && ((t >> (qp & 63)) & 1)) dv.add(stage[min(qp, 64u * kPackRounds - 1)]);
            }
        }
    }
}

// Dependencies of the entries settled in one bucket (P: IntPath or
// RealPath): a wavefront per entry walks its row 64 arcs at a time; for
// every arc whose head some lane of the entry reached, the lanes whose
// offer dist[v] + length ties the head's distance add its coefficient, in
// CSR order (up to kDialBackAhead rows loaded before they are added; node
// scores of entries with at most kPackLanes lanes: dial_sweep_packed). With
// kEdge every arc's credit also goes to eacc (edge_flush).
template <class C, bool kEdge, class P, bool kQuirk = false>
__global__ void __launch_bounds__(kBlock, NETBC_DIAL_BACK_WAVES)
    dial_backward_kernel(const QEnt* __restrict__ q, uint32_t qbeg, uint32_t qend,
                                     const uint32_t* __restrict__ off,
                                     const uint32_t* __restrict__ idx,
                                     const typename P::Len* __restrict__ len_arr,
                                     const typename P::Dist* __restrict__ dist,
                                     const uint64_t* __restrict__ vis, double* __restrict__ sig,
                                     C* __restrict__ coef, size_t n, double* __restrict__ eacc,
                                     size_t arcs, const uint8_t* __restrict__ target,
                                     uint32_t* __restrict__ npr, Ctrl* ctrl) {
    using Dist = typename P::Dist;
    using Len = typename P::Len;
    const int lane = lane_id();
    // entries by ticket (their rows differ widely in length)
    const uint32_t xc = blockIdx.x % kTicketClasses;
    // per wavefront: the packed coefficients of a window (dial_sweep_packed)
    __shared__ C stage_rows[kBlock / 64][64 * kPackRounds];
    C* stage = stage_rows[threadIdx.x >> 6];
    for (size_t e = qbeg + next_ticket(ctrl, xc); e < qend; e = qbeg + next_ticket(ctrl, xc)) {
        QEnt en = q[e];
        size_t base = (size_t)en.b * n;
        uint64_t M = en.m;
        const bool mine = (M >> lane) & 1;
        const double s = mine ? sig[(base + en.v) * kLanes + lane] : 0.0;
        const Dist dv0 = mine ? dist[(base + en.v) * kLanes + lane] : Dist(0);
        Dep<C> dv, dq;  // dq: kQuirk (as bfs_backward_kernel), non-target successors
        const uint64_t a0 = off[en.v], a1 = off[en.v + 1];
        bool packed = false;
        if constexpr (!kEdge && !kQuirk) {
            if ((uint32_t)__popcll(M) <= kPackLanes) {
                dial_sweep_packed<C, P>(a0, a1, M, base, dv0, idx, len_arr, dist, vis, coef, dv,
                                        stage);
                packed = true;
            }
        }
        for (uint64_t c = packed ? a1 : a0; c < a1; c += 64) {
            const uint64_t a = c + lane;
            bool ok = a < a1;
            uint32_t x = ok ? idx[a] : 0;
            Len len = ok ? len_arr[a] : Len(0);
            uint64_t mm = ok ? (vis[base + x] & M) : 0;
            uint64_t bal = __ballot(mm != 0);
            while (bal) {
                C val[kDialBackAhead];
                bool use[kDialBackAhead];
                uint32_t ak[kDialBackAhead], bk[kDialBackAhead];
                Dist dx[kDialBackAhead];
                Len lx[kDialBackAhead];
                size_t rx[kDialBackAhead];
                [[maybe_unused]] bool tk[kDialBackAhead];
                // the heads' distances first, then the coefficients of the
                // heads the arc leads to on a shortest path
#pragma unroll
                for (int u = 0; u < kDialBackAhead; u++) {
                    use[u] = false;
                    val[u] = 0;
                    ak[u] = kNone;
                    bk[u] = en.b;
                    tk[u] = true;
                    dx[u] = Dist(0);
                    lx[u] = Len(0);
                    rx[u] = 0;
                    if (bal) {
                        const int j = __ffsll((unsigned long long)bal) - 1;
                        bal &= bal - 1;
                        const uint32_t xx = rl32(x, j);
                        lx[u] = rl(len, j);
                        const uint64_t m2 = rl64(mm, j);
                        ak[u] = (uint32_t)(c + (uint32_t)j);
                        if constexpr (kQuirk) tk[u] = target_of(target, xx) != 0.0;
                        if ((m2 >> lane) & 1) {
                            rx[u] = (base + xx) * kLanes + lane;
                            use[u] = true;
                            dx[u] = dist[rx[u]];
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < kDialBackAhead; u++) {
                    use[u] = use[u] && P::tie(P::add(dv0, lx[u]), dx[u]);
                    if (use[u]) val[u] = coef[rx[u]];
                }
                [[maybe_unused]] double cv[kDialBackAhead];
#pragma unroll
                for (int u = 0; u < kDialBackAhead; u++) {
                    cv[u] = 0.0;
                    if (!use[u]) continue;
                    if (!kQuirk || tk[u]) {
                        dv.add(val[u]);
                        cv[u] = s * (double)val[u];
                    } else {
                        dq.add(val[u]);
                        cv[u] = (double)val[u];
                    }
                }
                if constexpr (kEdge) edge_flush<kDialBackAhead>(cv, ak, bk, eacc, arcs);
            }
        }
        if (mine) {
            size_t r = (base + en.v) * kLanes + lane;
            if constexpr (kQuirk) {
                const double de = s * dv.value() + dq.value();
                const uint32_t np = npr[r];
                npr[r] = 0;
                coef[r] = (C)(target_of(target, en.v) != 0.0 ? (1.0 + de) / s
                                                             : (np ? de / (double)np : 0.0));
                sig[r] = de;
            } else {
                double delta = s * dv.value();
                coef[r] = Dep<C>::coef(target_of(target, en.v), delta, s);
                sig[r] = delta;
            }
        }
    }
    retire_tickets(ctrl, xc);
}

// ---- integer lengths: the distance ring ------------------------------------
//
// The breadth-first engine carried over to integer lengths 1..W-1 (Dial's
// algorithm with exact integer distances, betweenness.rs:214-295): every
// integer distance D is a level. A ring of W bitmaps (slot D mod W) holds the
// lanes of each (slot, node) pair settled at the last W distances, and a ring
// of W offer bitmaps the lanes offered each of the next W - 1 distances. A
// step at distance D gathers (dial_gather_kernel): the pairs offered D in
// lanes not yet settled settle there, with path counts pulled over in-arcs
// from the lanes at distance D - len; then expands (dial_expand_kernel): the
// entries of distance D offer D + len along their out-arcs. Distances are
// never stored per lane: the level of a lane is the distance. Queue bounds
// per distance stay on the device, so steps chain without host waits; the
// host checks for the end every few steps. The dependency sweep is the
// breadth-first one (bfs_backward_kernel) with successors looked up at
// distance D + len (RingBits<false>), the ring rebuilt from the queue as it
// goes down.

// Settles distance D (d0 = D mod W). A wavefront takes 64 consecutive
// (slot, node) pairs of one slot (slots spread over the XCDs); a pair
// offered D takes N = offer & ~vis, clears the offer word, and, when N is
// not empty, gathers its path counts over its in-row (gather_counts with
// RingBits<true>, marking successor lanes in succ), joins vis and the
// distance-D bitmap with N and the level queue as (y, b, N). kNpr:
// predecessor-arc counts as well (npr).
template <bool kNpr>
__global__ void dial_gather_kernel(uint32_t D, uint32_t W, uint64_t* __restrict__ cand_ring,
                                   uint64_t* __restrict__ ring, size_t stride,
                                   const uint32_t* __restrict__ in_off,
                                   const uint32_t* __restrict__ in_idx,
                                   const uint32_t* __restrict__ in_len, uint64_t* __restrict__ vis,
                                   uint64_t* __restrict__ succ, double* __restrict__ sig,
                                   const uint8_t* __restrict__ heavy_flag, QEnt* q, uint32_t qcap,
                                   uint32_t* __restrict__ hq, Ctrl* ctrl, int slots, size_t n,
                                   uint32_t* __restrict__ npr) {
    const int lane = lane_id();
    const uint32_t d0 = D % W;
    uint64_t* cand = cand_ring + (size_t)d0 * stride;
    uint64_t* lev = ring + (size_t)d0 * stride;
    const RingBits<true> prev{ring, stride, in_len, d0, W};
    [[maybe_unused]] unsigned long long c4 = 0, c5 = 0;
    extern __shared__ unsigned long long pull_lds[];
    QueueStage stage{reinterpret_cast<QEnt*>(pull_lds) + (threadIdx.x >> 6) * kPullBuf, 0};
    const XcdWave xw = xcd_wave();
    const size_t gps = (n + 63) / 64;  // 64-pair groups per slot
    const size_t groups = (size_t)xcd_slots(xw.xcd, slots) * gps;
    for (size_t g = xw.w; g < groups; g += xw.stride) {
        const uint32_t b = xw.xcd + kXcds * (uint32_t)(g / gps);
        const uint32_t x0 = (uint32_t)(g % gps) * 64;
        const bool ok = x0 + (uint32_t)lane < n;
        const uint32_t x = ok ? x0 + lane : 0;
        const size_t p = (size_t)b * n + x;
        const uint64_t c = ok ? cand[p] : 0;
        if (!__ballot(c != 0)) continue;
        uint64_t V = 0, N = 0;
        uint32_t i0 = 0, deg = 0;
        if (c) {
            cand[p] = 0;  // the slot serves distance D + W next
            V = vis[p];
            N = c & ~V;
            if (N) {
                i0 = in_off[x];
                deg = in_off[x + 1] - i0;
            }
        }
        (void)gather_counts<false, kNpr>(x, b, N, i0, deg, in_idx, prev, succ, sig, n, c4, c5,
                                         npr);
        if (N) {
            vis[p] = V | N;
            lev[p] = N;
        }
        stage.add(N != 0, x, b, N, ctrl, q, qcap, hq, heavy_flag);
    }
    if (stage.cnt) stage.flush(ctrl, q, qcap, hq, heavy_flag);
    NETBC_FLUSH(ctrl, 10, c4);
    NETBC_FLUSH(ctrl, 11, c5);
}

// Offers from distance D (d0 = D mod W), once dial_gather_kernel settled it
// and heavy_scan_kernel cut its heavy rows: every entry (x, b, M) offers
// D + len to the lanes of each out-neighbour y not yet settled,
// nv = M & ~vis[y], ORed into the offer bitmap of distance D + len (len < W,
// so a slot of the ring other than d0). Light rows are taken up to 64
// entries per wavefront and walked as one concatenation, heavy rows as
// tasks of kChunkArcs arcs (hq / hpre, their count in ctrl->hchunks). Also
// records where distance D + 1 starts in the level queue and the heavy
// list, clears the ring slot of distance D + 1 of the entries of distance
// D + 1 - W (no longer looked back at), and raises ctrl->last_offer to the
// largest distance offered.
__global__ void dial_expand_kernel(uint32_t D, uint32_t W, const QEnt* __restrict__ q,
                                   uint32_t* __restrict__ lvl_q, uint32_t* __restrict__ lvl_h,
                                   const uint32_t* __restrict__ hq,
                                   const uint32_t* __restrict__ hpre,
                                   const uint32_t* __restrict__ off,
                                   const uint32_t* __restrict__ idx,
                                   const uint32_t* __restrict__ out_len,
                                   const uint64_t* __restrict__ vis,
                                   uint64_t* __restrict__ cand_ring, uint64_t* __restrict__ ring,
                                   size_t stride, const uint8_t* __restrict__ heavy_flag,
                                   Ctrl* ctrl, size_t n, uint32_t qcap) {
    const int lane = lane_id();
    const uint32_t d0 = D % W;
    // after an overflow the count runs past the queue (the host restarts)
    const uint32_t qend = min(ctrl->qlen, qcap);
    const uint32_t qbeg = min(lvl_q[D], qend);
    const uint32_t hbeg = lvl_h[D], hend = ctrl->hlen;
    const uint32_t hchunks = ctrl->hchunks;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    if (tid == 0) {
        lvl_q[D + 1] = qend;
        lvl_h[D + 1] = hend;
    }
    if (D + 1 >= W) {
        uint64_t* lev1 = ring + (size_t)((D + 1) % W) * stride;
        const uint32_t zb = lvl_q[D + 1 - W], ze = lvl_q[D + 2 - W];
        for (size_t i = zb + tid; i < ze; i += nth) {
            const QEnt e = q[i];
            lev1[(size_t)e.b * n + e.v] = 0;  // the slot holds only distance D + 1 - W
        }
    }
    uint32_t far = 0;  // largest distance this lane offered
    auto offer = [&](uint32_t y, uint32_t bb, uint64_t nv, uint32_t len) {
        if (!nv) return;
        uint32_t sl = d0 + len;
        if (sl >= W) sl -= W;
        unsigned long long* cp =
            (unsigned long long*)&cand_ring[(size_t)sl * stride + (size_t)bb * n + y];
        if ((*cp & nv) != nv) atomicOr(cp, (unsigned long long)nv);
        far = max(far, D + len);
    };
    const uint32_t nq = qend - qbeg;
    const uint32_t gsz = group_size(nq, wave_count());
    const size_t groups = ((size_t)nq + gsz - 1) / gsz;
    const size_t tasks = groups + hchunks;
    for (size_t w = global_wave(); w < tasks; w += wave_count()) {
        if (w < groups) {
            const size_t i = qbeg + w * gsz + lane;
            QEnt en{0, 0, 0};
            uint32_t a0 = 0, deg = 0;
            if ((uint32_t)lane < gsz && i < qend) {
                en = q[i];
                if (!heavy_flag[en.v]) {
                    a0 = off[en.v];
                    deg = off[en.v + 1] - a0;
                }
            }
            uint32_t T;
            const uint32_t pre = wave_prefix(deg, T);
            for (uint64_t t0 = 0; t0 < T; t0 += 64) {
                const uint32_t t = (uint32_t)t0 + lane;
                const int c = owner_lane(pre, t);
                const uint32_t bc = (uint32_t)__shfl((int)en.b, c, 64);
                const uint32_t ac = (uint32_t)__shfl((int)a0, c, 64);
                const uint32_t pc = (uint32_t)__shfl((int)pre, c, 64);
                const uint64_t mc = __shfl(en.m, c, 64);
                if (t < T) {
                    const uint32_t arc = ac + (t - pc);
                    const uint32_t y = idx[arc];
                    offer(y, bc, mc & ~vis[(size_t)bc * n + y], out_len[arc]);
                }
            }
        } else {
            // One chunk of a heavy row.
            const uint32_t t = (uint32_t)(w - groups);
            const uint32_t h = heavy_of_task(hpre, hbeg, hend, t);
            const QEnt en = q[hq[h]];
            const uint64_t a0 = (uint64_t)off[en.v] + (uint64_t)(t - hpre[h]) * kChunkArcs;
            const uint64_t a1 = min(a0 + kChunkArcs, (uint64_t)off[en.v + 1]);
            const size_t base = (size_t)en.b * n;
            for (uint64_t c = a0; c < a1; c += 64) {
                const uint64_t a = c + lane;
                if (a < a1) {
                    const uint32_t y = idx[a];
                    offer(y, en.b, en.m & ~vis[base + y], out_len[a]);
                }
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) far = max(far, (uint32_t)__shfl_xor((int)far, o, 64));
    if (lane == 0 && far > *(volatile uint32_t*)&ctrl->last_offer) atomicMax(&ctrl->last_offer, far);
}

// Clears the words of the queue entries [qb, qe) in `words` (one level's
// bitmap, whose words hold that level's lanes only).
__global__ void zero_entry_words_kernel(const QEnt* __restrict__ q, uint32_t qb, uint32_t qe,
                                        uint64_t* __restrict__ words, size_t n) {
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    for (size_t i = qb + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < qe; i += nth) {
        const QEnt e = q[i];
        words[(size_t)e.b * n + e.v] = 0;
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

// The reference's 4-ary indexed min-heap keyed by (distance, node)
// (weights.rs:155-277) over one thread's private arrays: keys hk, nodes hv
// and every node's heap position (kNone when absent).
struct Heap4 {
    double* hk;
    uint32_t* hv;
    uint32_t* pos;
    size_t len;

    __device__ void sift_up(size_t i) {
        const double kd = hk[i];
        const uint32_t kv = hv[i];
        while (i > 0) {
            const size_t p = (i - 1) / 4;
            if (!h_less(kd, kv, hk[p], hv[p])) break;
            hk[i] = hk[p];
            hv[i] = hv[p];
            pos[hv[i]] = (uint32_t)i;
            i = p;
        }
        hk[i] = kd;
        hv[i] = kv;
        pos[kv] = (uint32_t)i;
    }
    __device__ void sift_down(size_t i) {
        const double kd = hk[i];
        const uint32_t kv = hv[i];
        for (;;) {
            const size_t first = i * 4 + 1;
            if (first >= len) break;
            size_t best = first;
            const size_t end = first + 4 < len ? first + 4 : len;
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
    }
    __device__ void push_or_decrease(uint32_t node, double d) {
        const uint32_t p = pos[node];
        size_t i;
        if (p == kNone) {
            hk[len] = d;
            hv[len] = node;
            i = len++;
        } else {
            i = p;
            if (!h_less(d, node, hk[i], hv[i])) return;
            hk[i] = d;
        }
        sift_up(i);
    }
    // Removes the least (distance, node); false when the heap is empty.
    __device__ bool pop(double& d, uint32_t& v) {
        if (len == 0) return false;
        d = hk[0];
        v = hv[0];
        pos[v] = kNone;
        len--;
        if (len > 0) {
            hk[0] = hk[len];
            hv[0] = hv[len];
            sift_down(0);
        }
        return true;
    }
};

// Dijkstra with the reference's 4-ary indexed heap ordered by (distance,
// node) (weights.rs:155-277) and ties within a relative 1e-10
// (betweenness.rs:117-120, 297-368). Each thread owns private arrays. With
// eacc, thread t adds every arc's credit to its own row eacc[t * arcs ..].
__global__ void dijkstra_kernel(const int64_t* __restrict__ list, int64_t item0, int64_t k,
                                int slots, size_t n, const uint32_t* __restrict__ off,
                                const uint32_t* __restrict__ idx, const double* __restrict__ lens,
                                DjkState st, uint64_t* __restrict__ vis, double* __restrict__ sig,
                                Ctrl* ctrl, double* __restrict__ eacc, size_t arcs,
                                const uint8_t* __restrict__ target) {
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
    uint32_t* pos = st.pos + o;
    for (size_t v = 0; v < n; v++) {
        dist[v] = INFINITY;
        sigma[v] = 0.0;
        coef[v] = 0.0;
        rank[v] = kNone;
        pos[v] = kNone;
    }
    size_t norder = 0;
    Heap4 heap{st.hkey + o, st.hnode + o, pos, 0};
    auto approx_eq = [](double a, double c) {
        return fabs(a - c) <= 1e-10 * fmax(fabs(a), fabs(c));
    };
    dist[s] = 0.0;
    sigma[s] = 1.0;
    heap.push_or_decrease(s, 0.0);
    double d;
    uint32_t v;
    while (heap.pop(d, v)) {
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
                heap.push_or_decrease(w, alt);
            } else if (approx_eq(alt, cur)) {
                sigma[w] += sigma[v];
            } else if (alt < cur) {
                dist[w] = alt;
                sigma[w] = sigma[v];
                heap.push_or_decrease(w, alt);
            }
        }
    }
    double mx = 0.0;
    for (size_t i = norder; i-- > 0;) {
        v = order[i];
        uint32_t rv = rank[v];
        double dvv = dist[v];
        double dv = 0.0;
        double* erow = eacc ? eacc + (size_t)t * arcs : nullptr;
        for (uint32_t a = off[v]; a < off[v + 1]; a++) {
            uint32_t w = idx[a];
            if (rank[w] > rv && rank[w] != kNone && approx_eq(dvv + lens[a], dist[w])) {
                const double cw = coef[w];
                dv += cw;
                if (erow) erow[a] += sigma[v] * cw;
            }
        }
        double delta = sigma[v] * dv;
        coef[v] = (target_of(target, v) + delta) / sigma[v];
        size_t mi = (size_t)b * n + v;
        atomicOr((unsigned long long*)&vis[mi], 1ull << l);
        sig[mi * kLanes + l] = delta;
        mx = fmax(mx, sigma[v]);
    }
    atomicMax(&ctrl->max_sigma_bits, (unsigned long long)__double_as_longlong(mx));
    if (mx > kTwo53) ctrl->sig_overflow = 1;
}

// ---- batch composition ----------------------------------------------------

// Runs with at least this many sources batch them by neighbourhood; smaller
// runs keep list order (and so the reference's summation order).
constexpr int64_t kReorderMin = 4096;

// Landmark distances for batching sources: search along reversed arcs from
// each landmark (lane k of a 32-bit mask = landmark k), level by level up
// to level 15. landmark_offer_kernel: every frontier node v (cur[v] != 0)
// offers its lanes to its in-neighbours u not yet reached in them (a
// wavefront per frontier node, 64 in-arcs at a time);
// landmark_settle_kernel: new lanes become reached and form the next
// frontier, and dist[u][k] = level for them.
__global__ void landmark_offer_kernel(const uint32_t* __restrict__ in_off,
                                      const uint32_t* __restrict__ in_idx,
                                      const uint32_t* __restrict__ cur,
                                      const uint32_t* __restrict__ vis,
                                      uint32_t* __restrict__ nxt, size_t n) {
    const int lane = lane_id();
    for (size_t v0 = global_wave() * 64; v0 < n; v0 += wave_count() * 64) {
        const size_t v = v0 + lane;
        const uint32_t M = v < n ? cur[v] : 0u;
        uint64_t act = __ballot(M != 0);
        while (act) {
            const int j = __ffsll((unsigned long long)act) - 1;
            act &= act - 1;
            const uint32_t vj = (uint32_t)(v0 + j);
            const uint32_t Mj = rl32(M, j);
            const uint32_t a1 = in_off[vj + 1];
            for (uint64_t a = (uint64_t)in_off[vj] + lane; a < a1; a += 64) {
                const uint32_t u = in_idx[a];
                const uint32_t nw = Mj & ~vis[u];
                if (nw) atomicOr(&nxt[u], nw);
            }
        }
    }
}

__global__ void landmark_settle_kernel(uint32_t* __restrict__ cur, uint32_t* __restrict__ vis,
                                       uint32_t* __restrict__ nxt, uint8_t* __restrict__ dist,
                                       int level, int lanes, size_t n) {
    const size_t nth = (size_t)gridDim.x * blockDim.x;
    for (size_t u = (size_t)blockIdx.x * blockDim.x + threadIdx.x; u < n; u += nth) {
        const uint32_t N = nxt[u] & ~vis[u];
        nxt[u] = 0;
        cur[u] = N;
        if (!N) continue;
        vis[u] |= N;
        for (uint32_t M = N; M; M &= M - 1) {
            const int k = __ffs((int)M) - 1;
            if (k < lanes) dist[u * kLandmarks + k] = (uint8_t)level;
        }
    }
}

// Sort key of the source at position i of the current order perm (which
// 0: second out-neighbour, 1: first out-neighbour, 2: landmark distances,
// landmark 0 in the top four bits). Sorting stably by the three in turn
// batches sources whose searches look alike: same distances to the
// landmarks, then same first neighbours.
// Landmarks whose weighted distances key the batching of integer-length
// sources (a byte each in a 64-bit key).
constexpr int kWeightedLandmarks = 8;

__global__ void source_key_kernel(const int64_t* __restrict__ list, int64_t k,
                                  const uint32_t* __restrict__ perm,
                                  const uint32_t* __restrict__ off,
                                  const uint32_t* __restrict__ idx,
                                  const uint8_t* __restrict__ dist, uint32_t n, int which,
                                  uint64_t* __restrict__ key) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    const uint32_t p = perm[i];
    const uint32_t s = (uint32_t)(list ? list[p] : p);
    const uint32_t a0 = off[s], a1 = off[s + 1];
    uint64_t kv;
    if (which == 0) {
        kv = a1 > a0 + 1 ? idx[a0 + 1] : n;
    } else if (which == 1) {
        kv = a1 > a0 ? idx[a0] : n;
    } else if (which == 2) {
        kv = 0;
        const uint8_t* d = dist + (size_t)s * kLandmarks;
        for (int j = 0; j < kLandmarks; j++) kv = (kv << 4) | (uint64_t)(d[j] & 15u);
    } else {
        // weighted distances to the first kWeightedLandmarks landmarks, a byte each
        kv = 0;
        const uint8_t* d = dist + (size_t)s * kLandmarks;
        for (int j = 0; j < kWeightedLandmarks; j++) kv = (kv << 8) | (uint64_t)d[j];
    }
    key[i] = kv;
}

// One Bellman-Ford round for the weighted distances from every node to the
// landmarks (integer lengths): every node v takes, per landmark j <
// kWeightedLandmarks, the least of wd[v][j] and len(v, u) + wd[u][j] over
// its out-neighbours u (a wavefront per node, lanes striding its row, then
// a wavefront minimum). Updates in place (the minima only fall, so the
// rounds converge to the distances); *changed is set if any fell.
// (Real lengths: FP64 distances, INFINITY when unreached.)
template <class T>
__global__ void landmark_relax_kernel(const uint32_t* __restrict__ off,
                                      const uint32_t* __restrict__ idx,
                                      const T* __restrict__ len, T* __restrict__ wd,
                                      size_t n, uint32_t* __restrict__ changed) {
    constexpr T kFar = std::is_same_v<T, uint32_t> ? (T)kNone : (T)INFINITY;
    const int lane = lane_id();
    for (size_t v = global_wave(); v < n; v += wave_count()) {
        T best[kWeightedLandmarks];
#pragma unroll
        for (int j = 0; j < kWeightedLandmarks; j++) best[j] = kFar;
        const uint64_t a1 = off[v + 1];
        for (uint64_t a = (uint64_t)off[v] + lane; a < a1; a += 64) {
            const T* du = wd + (size_t)idx[a] * kWeightedLandmarks;
            const T l = len[a];
#pragma unroll
            for (int j = 0; j < kWeightedLandmarks; j++)
                if (du[j] != kFar) best[j] = min(best[j], du[j] + l);
        }
        bool fell = false;
#pragma unroll
        for (int j = 0; j < kWeightedLandmarks; j++) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) best[j] = min(best[j], __shfl_xor(best[j], o, 64));
            const T old = wd[v * kWeightedLandmarks + j];
            if (best[j] < old) {
                if (lane == 0) wd[v * kWeightedLandmarks + j] = best[j];
                fell = true;
            }
        }
        if (lane == 0 && fell) *changed = 1;
    }
}

// min(wd / unit, 255) as bytes into the landmark distance table (first
// kWeightedLandmarks columns of each node's kLandmarks); unit 1 for integer
// lengths, the shortest arc for real ones.
template <class T>
__global__ void landmark_bytes_kernel(const T* __restrict__ wd, size_t n, double unit,
                                      uint8_t* __restrict__ dist) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        for (int j = 0; j < kWeightedLandmarks; j++) {
            const double q = (double)wd[v * kWeightedLandmarks + j] / unit;
            dist[v * kLandmarks + j] = (uint8_t)(q < 255.0 ? q : 255.0);
        }
}

__global__ void iota_kernel(uint32_t* __restrict__ v, int64_t k) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) v[i] = (uint32_t)i;
}

__global__ void gather_sources_kernel(const int64_t* __restrict__ list, int64_t k,
                                      const uint32_t* __restrict__ perm,
                                      int64_t* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < k) out[i] = list ? list[perm[i]] : (int64_t)perm[i];
}

// ---- cross-source reduction -----------------------------------------------

// Adds this super-batch's dependencies into the per-node group
// accumulator in item order. When an item belongs to a later group than
// the last one added, the accumulator is first merged into bc and reset
// (groups without contributions would add +0.0, which changes nothing).
// This is sum_in_order's sequence of operations on every node.
__global__ void accumulate_kernel(const uint64_t* __restrict__ vis, double* __restrict__ sig,
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
            double* row = sig + ((size_t)b * n + v) * kLanes;
            while (M) {
                int l = __ffsll((unsigned long long)M) - 1;
                M &= M - 1;
                int64_t i = ib + l;
                int64_t s = list ? list[i] : i;
                double x = row[l];
                row[l] = 0.0;  // rows start at zero for the next super-batch
                if ((size_t)s == v) continue;
                int64_t gi = i / gsize;
                if (gi != g) {
                    B += a;
                    a = 0.0;
                    g = gi;
                }
                a += x;
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
        atomicAdd(&ctrl->reached_nodes[blockIdx.x % kSpread], rn);
        atomicAdd(&ctrl->reached_arcs[blockIdx.x % kSpread], ra);
    }
}

// The same sums with a wavefront per node, for super-batches that span
// many groups (few sources per group): lane j takes group gA + j (gA, gB:
// the first and last group of this super-batch's items), adds its items'
// dependencies in item order from 0.0 (group gA from the carried partial
// sum when it continues one), and the group sums are then added into bc in
// group order; the last group's sum is carried. Groups are summed in the
// same order and each group's items in the same order as accumulate_kernel
// (a group without contributions adds +0.0, which changes nothing), so the
// bits are the same.
__global__ void accumulate_groups_kernel(const uint64_t* __restrict__ vis,
                                         double* __restrict__ sig, int slots, size_t n,
                                         int64_t item0, int64_t k, int64_t gsize,
                                         const int64_t* __restrict__ list,
                                         double* __restrict__ acc, double* __restrict__ bc,
                                         int64_t* __restrict__ gcur,
                                         const uint32_t* __restrict__ off, Ctrl* ctrl) {
    const int lane = lane_id();
    const int64_t iend = min(k, item0 + (int64_t)slots * kLanes);
    const int64_t gA = item0 / gsize, gB = (iend - 1) / gsize;
    const int active = (int)((iend - item0 + kLanes - 1) / kLanes);  // slots with items
    unsigned long long rn = 0, ra = 0;
    for (size_t v = global_wave(); v < n; v += wave_count()) {
        double a = acc[v], B = bc[v];
        const int64_t g = gcur[v];
        if (g < gA) {  // the carried group ends before this super-batch
            B += a;
            a = 0.0;
        }
        for (int64_t g0 = gA; g0 <= gB; g0 += kLanes) {
            const int64_t gj = g0 + lane;
            double sum = gj == g ? a : 0.0;
            if (gj <= gB) {
                const int64_t i0 = max(gj * gsize, item0), i1 = min((gj + 1) * gsize, iend);
                for (int64_t i = i0; i < i1; i++) {
                    const int64_t rel = i - item0;
                    const size_t p = (size_t)(rel >> 6) * n + v;
                    const int l = (int)(rel & 63);
                    if ((vis[p] >> l) & 1) {
                        double* x = &sig[p * kLanes + l];
                        const double d = *x;
                        *x = 0.0;  // rows start at zero for the next super-batch
                        const int64_t src = list ? list[i] : i;
                        if ((size_t)src != v) sum += d;
                    }
                }
            }
            for (int j = 0; j < kLanes && g0 + j <= gB; j++) {
                const double sj = __shfl(sum, j, 64);
                if (g0 + j < gB) B += sj;
                else a = sj;
            }
        }
        unsigned long long r = 0;
        for (int b = lane; b < active; b += kLanes)
            r += (unsigned long long)__popcll(vis[(size_t)b * n + v]);
        const unsigned long long rv = wave_sum_u64(r);
        if (lane == 0) {
            acc[v] = a;
            bc[v] = B;
            gcur[v] = gB;
            rn += rv;
            ra += rv * (unsigned long long)(off[v + 1] - off[v]);
        }
    }
    if (lane == 0 && (rn || ra)) {
        atomicAdd(&ctrl->reached_nodes[blockIdx.x % kSpread], rn);
        atomicAdd(&ctrl->reached_arcs[blockIdx.x % kSpread], ra);
    }
}

// Zeroes the per-lane rows of every reached (node, lane) of the slots, so
// the path-count rows start at zero for the next super-batch.
__global__ void clear_rows_kernel(const uint64_t* __restrict__ vis, double* __restrict__ sig,
                                  size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint64_t M = vis[i];
    while (M) {
        int l = __ffsll((unsigned long long)M) - 1;
        M &= M - 1;
        sig[i * kLanes + l] = 0.0;
    }
}

// Zeroes the reached lanes of every hub's row (bfs_hub_kernel adds hub
// path counts atomically, so those rows must start at zero): a wavefront
// per (slot, hub) row, lane l clearing entry l if l reached the hub.
__global__ void zero_hub_rows_kernel(const uint64_t* __restrict__ vis, double* __restrict__ sig,
                                     const uint32_t* __restrict__ hub_list, uint32_t hubs,
                                     int slots, size_t n) {
    const int lane = lane_id();
    const size_t rows = (size_t)slots * hubs;
    for (size_t r = global_wave(); r < rows; r += wave_count()) {
        const size_t p = (r / hubs) * n + hub_list[r % hubs];
        if ((vis[p] >> lane) & 1) sig[p * kLanes + lane] = 0.0;
    }
}

// Zeroes each source's own entry (sigma = 1 at level 0, which the sweep
// leaves in place) so that the row sums below need no per-lane source test.
__global__ void zero_sources_kernel(const QEnt* __restrict__ q, uint32_t count,
                                    double* __restrict__ sig, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const QEnt e = q[i];
    uint64_t M = e.m;
    while (M) {
        int l = __ffsll((unsigned long long)M) - 1;
        M &= M - 1;
        sig[((size_t)e.b * n + e.v) * kLanes + l] = 0.0;
    }
}

// Scores of a free-order breadth-first super-batch: a thread per node v
// adds bsum[b][v] (the sweep's dependency sums) over the active slots b in
// slot order to bc[v] (deterministic for a given slot count) and clears
// bsum for the next super-batch; neighbouring threads read neighbouring
// words of each slot's row, 8 slots' words in flight at a time. Also
// counts the reached (source, node) pairs and their out-arcs. Sources' own
// entries (level 0) add nothing: the sweep adds no sums at level 0.
__global__ void fold_slot_sums_kernel(const uint64_t* __restrict__ vis, double* __restrict__ bsum,
                                      int active, size_t n, double* __restrict__ bc,
                                      const uint32_t* __restrict__ off, Ctrl* ctrl) {
    constexpr int kBatch = 8;
    unsigned long long rn = 0, ra = 0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        double x = 0.0;
        unsigned long long cnt = 0;
        for (int b0 = 0; b0 < active; b0 += kBatch) {
            uint64_t M[kBatch];
            double y[kBatch];
#pragma unroll
            for (int j = 0; j < kBatch; j++)
                M[j] = b0 + j < active ? vis[(size_t)(b0 + j) * n + v] : 0;
#pragma unroll
            for (int j = 0; j < kBatch; j++) y[j] = M[j] ? bsum[(size_t)(b0 + j) * n + v] : 0.0;
#pragma unroll
            for (int j = 0; j < kBatch; j++) {
                cnt += (unsigned long long)__popcll(M[j]);
                if (y[j] != 0.0) {
                    x += y[j];
                    bsum[(size_t)(b0 + j) * n + v] = 0.0;
                }
            }
        }
        if (x != 0.0) bc[v] += x;
        rn += cnt;
        ra += cnt * (unsigned long long)(off[v + 1] - off[v]);
    }
    rn = wave_sum_u64(rn);
    ra = wave_sum_u64(ra);
    if (lane_id() == 0 && (rn || ra)) {
        atomicAdd(&ctrl->reached_nodes[blockIdx.x % kSpread], rn);
        atomicAdd(&ctrl->reached_arcs[blockIdx.x % kSpread], ra);
    }
}

// The same sum for runs whose source order is free (neighbourhood
// batching). A wave per node v: lane b reads the reached lanes of v in
// slot b (64 slots at a time), then the rows of the slots that reached v
// are read kAhead at a time and added per lane in slot order; one fixed
// butterfly adds the lanes, and bc[v] += the result (deterministic for a
// given slot count). Sources' own entries are zero (zero_sources_kernel).
// The rows are left as they are: the next super-batch's search stores every
// count it reads before reading it, except at hubs, whose counts are added
// atomically and whose rows zero_hub_rows_kernel clears.
__global__ void accumulate_rows_kernel(const uint64_t* __restrict__ vis, double* __restrict__ sig,
                                       int slots, size_t n, int64_t item0, int64_t k,
                                       double* __restrict__ bc,
                                       const uint32_t* __restrict__ off, Ctrl* ctrl) {
    const int lane = lane_id();
    unsigned long long rn = 0, ra = 0;
    for (size_t v = global_wave(); v < n; v += wave_count()) {
        double x = 0.0;
        unsigned long long cnt = 0;
        for (int b0 = 0; b0 < slots; b0 += 64) {
            const int b = b0 + lane;
            uint64_t M = 0;
            if (b < slots && item0 + (int64_t)b * kLanes < k) M = vis[(size_t)b * n + v];
            cnt += (unsigned long long)__popcll(M);
            uint64_t act = __ballot(M != 0);
            while (act) {
                double val[kAhead];
#pragma unroll
                for (int j = 0; j < kAhead; j++) {
                    val[j] = 0.0;
                    if (act) {
                        const int s = __ffsll((unsigned long long)act) - 1;
                        act &= act - 1;
                        const uint64_t Ms = rl64(M, s);
                        if ((Ms >> lane) & 1) {
                            double* r = &sig[((size_t)(b0 + s) * n + v) * kLanes + lane];
                            val[j] = *r;
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < kAhead; j++) x += val[j];
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        cnt = wave_sum_u64(cnt);
        if (lane == 0) {
            if (x != 0.0) bc[v] += x;
            rn += cnt;
            ra += cnt * (unsigned long long)(off[v + 1] - off[v]);
        }
    }
    rn = wave_sum_u64(rn);
    ra = wave_sum_u64(ra);
    if (lane == 0 && (rn || ra)) {
        atomicAdd(&ctrl->reached_nodes[blockIdx.x % kSpread], rn);
        atomicAdd(&ctrl->reached_arcs[blockIdx.x % kSpread], ra);
    }
}

// Final merge of the last group.
__global__ void finish_kernel(double* __restrict__ bc, const double* __restrict__ acc, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        bc[v] += acc[v];
}

__global__ void scale_kernel(double* __restrict__ bc, size_t n, double scale) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        bc[v] *= scale;
}

// Edge sums: every arc's credit summed over the accumulator rows in row
// order, then added to its input edge. An edge has at most two arcs and
// starts at zero, and two-term sums do not depend on their order, so the
// atomic adds leave a deterministic result.
__global__ void fold_arcs_kernel(const double* __restrict__ eacc, int rows, size_t arcs,
                                 const uint32_t* __restrict__ arc_edge, double* __restrict__ eb) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < arcs; k += stride) {
        double x = 0.0;
        for (int r = 0; r < rows; r++) x += eacc[(size_t)r * arcs + k];
        if (x != 0.0) atomicAdd(&eb[arc_edge[k]], x);
    }
}

// Path endpoints (path_centrality.rs:480-492): every target t reached from
// a source s != t adds 1 to t and 1 to s. endpoint_nodes_kernel adds to
// ep[v], for a target v, the lanes of the super-batch that reached it
// (including a lane whose source is v); endpoint_lanes_kernel counts, per
// (slot, lane), the targets the lane reached (including its source when
// that is a target), 64 nodes per load: lane l tests bit l of each nonzero
// loaded word in turn; a wavefront keeps to one slot and adds its counts
// once. endpoint_sources_kernel adds each lane's count to
// its source, less 2 when the source is a target (the two self counts).
// The counts are integers, so the order of the atomic adds does not matter.
__global__ void endpoint_nodes_kernel(const uint64_t* __restrict__ vis, int slots, size_t n,
                                      const uint8_t* __restrict__ target,
                                      unsigned long long* __restrict__ ep) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        if (target && !target[v]) continue;
        unsigned long long c = 0;
        for (int b = 0; b < slots; b++) c += (unsigned long long)__popcll(vis[(size_t)b * n + v]);
        if (c) ep[v] += c;
    }
}

__global__ void endpoint_lanes_kernel(const uint64_t* __restrict__ vis, int slots, size_t n,
                                      const uint8_t* __restrict__ target,
                                      unsigned long long* __restrict__ lane_count) {
    const int lane = lane_id();
    const size_t chunks = (n + 63) / 64;
    // wps wavefronts per slot, each taking every wps-th chunk of its slot and
    // adding its lane counts once at the end
    const size_t G = wave_count();
    const size_t wps = std::max<size_t>(1, G / (size_t)slots);
    for (size_t idx = global_wave(); idx < (size_t)slots * wps; idx += G) {
        const size_t b = idx / wps;
        unsigned long long c = 0;
        for (size_t ch = idx % wps; ch < chunks; ch += wps) {
            const size_t v = ch * 64 + lane;
            uint64_t word = 0;
            if (v < n && (!target || target[v])) word = vis[b * n + v];
            uint64_t nz = __ballot(word != 0);  // only the nonzero words
            while (nz) {
                const int j = __ffsll((unsigned long long)nz) - 1;
                nz &= nz - 1;
                c += (rl64(word, j) >> lane) & 1;
            }
        }
        if (c) atomicAdd(&lane_count[b * kLanes + lane], c);
    }
}

__global__ void endpoint_sources_kernel(const int64_t* __restrict__ list, int64_t item0,
                                        int64_t k, int slots, const uint8_t* __restrict__ target,
                                        unsigned long long* __restrict__ lane_count,
                                        unsigned long long* __restrict__ ep) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= slots * kLanes) return;
    const int64_t i = item0 + t;
    if (i < k) {
        const int64_t s = list ? list[i] : i;
        const unsigned long long self = (!target || target[s]) ? 2ull : 0ull;
        atomicAdd(&ep[s], lane_count[t] - self);  // modular: the sum stays exact
    }
    lane_count[t] = 0;
}

__global__ void add_counts_kernel(double* __restrict__ bc, const unsigned long long* __restrict__ ep,
                                  size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        if (ep[v]) bc[v] += (double)ep[v];
}

// ---- distance profiles (closeness.rs:212-223, 240-345) --------------------

// Per source: other nodes reached, the sum of their lengths (an integer for
// hop counts and integer lengths, else a double), the sum of reciprocal
// lengths and the longest length. Five arrays of `count` 8-byte words, one
// after another from `base` (all zero bits: an empty profile).
struct DevProfile {
    unsigned long long* reached;
    unsigned long long* itotal;
    double* ftotal;
    double* harmonic;
    double* longest;
};

inline DevProfile profile_view(uint64_t* base, size_t count) {
    return DevProfile{(unsigned long long*)base, (unsigned long long*)base + count,
                      (double*)(base + 2 * count), (double*)(base + 3 * count),
                      (double*)(base + 4 * count)};
}

// Breadth-first level `level` (per-lane counts from endpoint_lanes_kernel
// over the level's bitmap): lane t's profile gets reached += c, total += c *
// level, harmonic += c / level and longest = level, in level order as
// MultiBfs's per-level bookkeeping (closeness.rs:268-293). The counts are
// reset for the next level.
__global__ void level_profile_kernel(unsigned long long* __restrict__ lane_count, int slots,
                                     uint32_t level, DevProfile lanes) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= slots * kLanes) return;
    const unsigned long long c = lane_count[t];
    if (c == 0) return;
    lane_count[t] = 0;
    lanes.reached[t] += c;
    lanes.itotal[t] += c * level;
    lanes.harmonic[t] += (double)c / (double)level;
    lanes.longest[t] = (double)level;
}

// Bucketed searches: partial profiles per (slot, chunk of kProfileChunk
// nodes, lane) from the final distance rows, nodes in order; the source
// itself (distance 0) is left out.
constexpr uint32_t kProfileChunk = 512;

template <class P>
__global__ void dist_profile_kernel(const typename P::Dist* __restrict__ dist, int slots,
                                    size_t n, uint32_t* __restrict__ pc,
                                    unsigned long long* __restrict__ pit,
                                    double* __restrict__ pft, double* __restrict__ ph,
                                    double* __restrict__ pl) {
    using Dist = typename P::Dist;
    const int lane = lane_id();
    const size_t chunks = (n + kProfileChunk - 1) / kProfileChunk;
    for (size_t w = global_wave(); w < chunks * (size_t)slots; w += wave_count()) {
        const size_t b = w / chunks, ch = w % chunks;
        const size_t v0 = ch * kProfileChunk, v1 = min(n, v0 + kProfileChunk);
        uint32_t cnt = 0;
        unsigned long long it = 0;
        double ft = 0.0, h = 0.0, lg = 0.0;
        for (size_t v = v0; v < v1; v++) {
            const Dist d = dist[(b * n + v) * kLanes + lane];
            if (!P::reached(d) || d == Dist(0)) continue;
            cnt++;
            if constexpr (std::is_same_v<Dist, uint32_t>) it += d;
            else ft += d;
            h += 1.0 / (double)d;
            lg = fmax(lg, (double)d);
        }
        const size_t r = w * kLanes + lane;
        pc[r] = cnt;
        pit[r] = it;
        pft[r] = ft;
        ph[r] = h;
        pl[r] = lg;
    }
}

// The partial profiles of each (slot, lane) added in chunk order.
__global__ void dist_profile_reduce_kernel(const uint32_t* __restrict__ pc,
                                           const unsigned long long* __restrict__ pit,
                                           const double* __restrict__ pft,
                                           const double* __restrict__ ph,
                                           const double* __restrict__ pl, int slots, size_t n,
                                           DevProfile lanes) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= slots * kLanes) return;
    const int b = t / kLanes, l = t % kLanes;
    const size_t chunks = (n + kProfileChunk - 1) / kProfileChunk;
    unsigned long long c = 0, it = 0;
    double ft = 0.0, h = 0.0, lg = 0.0;
    for (size_t ch = 0; ch < chunks; ch++) {
        const size_t r = ((size_t)b * chunks + ch) * kLanes + l;
        c += pc[r];
        it += pit[r];
        ft += pft[r];
        h += ph[r];
        lg = fmax(lg, pl[r]);
    }
    lanes.reached[t] = c;
    lanes.itotal[t] = it;
    lanes.ftotal[t] = ft;
    lanes.harmonic[t] = h;
    lanes.longest[t] = lg;
}

// The lane profiles of a super-batch, written to their sources.
__global__ void commit_profile_kernel(DevProfile lanes, int slots, const int64_t* __restrict__ list,
                                      int64_t item0, int64_t k, DevProfile prof) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= slots * kLanes) return;
    const int64_t i = item0 + t;
    if (i >= k) return;
    const int64_t s = list ? list[i] : i;
    prof.reached[s] = lanes.reached[t];
    prof.itotal[s] = lanes.itotal[t];
    prof.ftotal[s] = lanes.ftotal[t];
    prof.harmonic[s] = lanes.harmonic[t];
    prof.longest[s] = lanes.longest[t];
}

// Real lengths with very unequal arcs: the reference's Dijkstra
// (closeness.rs:427-470), one source per thread on private arrays (st.dist,
// st.rank as the settled mark, the heap in st.hkey / st.hnode / st.pos).
// Every settled node other than the source adds to lane t's profile in
// settle order, as the reference does.
__global__ void dijkstra_profile_kernel(const int64_t* __restrict__ list, int64_t item0,
                                        int64_t k, int slots, size_t n,
                                        const uint32_t* __restrict__ off,
                                        const uint32_t* __restrict__ idx,
                                        const double* __restrict__ lens, DjkState st,
                                        DevProfile lanes) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= slots * kLanes) return;
    const int64_t item = item0 + t;
    if (item >= k) return;
    const uint32_t s = (uint32_t)(list ? list[item] : item);
    const size_t o = (size_t)t * n;
    double* dist = st.dist + o;
    uint32_t* done = st.rank + o;
    uint32_t* pos = st.pos + o;
    for (size_t v = 0; v < n; v++) {
        dist[v] = INFINITY;
        done[v] = 0;
        pos[v] = kNone;
    }
    Heap4 heap{st.hkey + o, st.hnode + o, pos, 0};
    dist[s] = 0.0;
    heap.push_or_decrease(s, 0.0);
    unsigned long long reached = 0;
    double total = 0.0, harmonic = 0.0, longest = 0.0;
    double d;
    uint32_t v;
    while (heap.pop(d, v)) {
        done[v] = 1;
        if (v != s) {
            reached++;
            total += d;
            harmonic += 1.0 / d;
            if (d > longest) longest = d;
        }
        for (uint32_t a = off[v]; a < off[v + 1]; a++) {
            const uint32_t w = idx[a];
            if (done[w]) continue;
            const double alt = d + lens[a];
            if (alt < dist[w]) {
                dist[w] = alt;
                heap.push_or_decrease(w, alt);
            }
        }
    }
    lanes.reached[t] = reached;
    lanes.itotal[t] = 0;
    lanes.ftotal[t] = total;
    lanes.harmonic[t] = harmonic;
    lanes.longest[t] = longest;
}

// Closeness (harmonic == 0) or harmonic centrality from the profiles,
// closeness.rs:131-209 (unit: the common length of uniform weights, which
// hop counts are multiplied by). rule: 0 ignore, 1 Wasserman-Faust, 2 zero,
// 3 max plus one (longest: the largest finite length of any source).
__global__ void closeness_kernel(DevProfile prof, size_t n, int use_ftotal, double unit,
                                 int harmonic, int normalized, int rule,
                                 const unsigned long long* __restrict__ longest_bits,
                                 double* __restrict__ out) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        double total = use_ftotal ? prof.ftotal[v] : (double)prof.itotal[v];
        double h = prof.harmonic[v];
        if (unit != 1.0) {
            total *= unit;
            h /= unit;
        }
        if (harmonic) {
            out[v] = h * (normalized && n > 1 ? 1.0 / (double)(n - 1) : 1.0);
            continue;
        }
        const unsigned long long reached = prof.reached[v];
        const double r = (double)reached, others = n > 0 ? (double)(n - 1) : 0.0;
        double x;
        if (rule == 0) {
            x = reached == 0 ? NAN : (normalized ? r / total : 1.0 / total);
        } else if (rule == 1) {
            x = reached == 0 ? 0.0 : (r / total) * (r / others);
        } else if (rule == 2) {
            x = n < 2 ? NAN
                : reached < n - 1 ? 0.0
                : normalized ? others / total : 1.0 / total;
        } else {
            double longest = __longlong_as_double((long long)*longest_bits);
            if (unit != 1.0) longest *= unit;
            const double t2 = total + (others - r) * (longest + 1.0);
            x = n < 2 ? NAN : normalized ? others / t2 : 1.0 / t2;
        }
        out[v] = x;
    }
}

// Closeness normalized for two-mode data (bipartite.rs:176-210): with own
// and other the sizes of v's mode and of the other mode, (other + 2(own -
// 1)) / total when normalized, else 1 / total; times reached / (n - 1)
// under the Wasserman-Faust rule (wf), which gives 0 to a node reaching no
// other (NaN otherwise).
__global__ void twomode_closeness_kernel(DevProfile prof, size_t n,
                                         const uint8_t* __restrict__ types, double first,
                                         double second, int normalized, int wf,
                                         double* __restrict__ out) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const unsigned long long r = prof.reached[v];
        if (r == 0) {
            out[v] = wf ? 0.0 : NAN;
            continue;
        }
        const double total = (double)prof.itotal[v];
        const double own = types[v] ? second : first, other = types[v] ? first : second;
        double c = normalized ? (other + 2.0 * (own - 1.0)) / total : 1.0 / total;
        if (wf) c *= (double)r / (double)(n - 1);
        out[v] = c;
    }
}

// Degree of every node from the edge list, degree.rs:41-74 (mode 1 out, 2
// in, 3 all; loops skipped unless counted). Integer counts: the order of
// the atomic adds does not matter.
__global__ void degree_kernel(const uint32_t* __restrict__ src, const uint32_t* __restrict__ tgt,
                              size_t m, int directed, int mode, int loops,
                              double* __restrict__ out) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
        const uint32_t s = src[e], t = tgt[e];
        if (s == t && !loops) continue;
        if (!directed || mode != 2) atomicAdd(&out[s], 1.0);
        if (!directed || mode != 1) atomicAdd(&out[t], 1.0);
    }
}

// vals[v] /= div[types[v]] where that divisor is positive, else NaN.
__global__ void divide_by_type_kernel(double* __restrict__ vals, size_t n,
                                      const uint8_t* __restrict__ types, double div0,
                                      double div1) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const double d = types[v] ? div1 : div0;
        vals[v] = d > 0.0 ? vals[v] / d : NAN;
    }
}

// The largest longest length, as bits (non-negative doubles order as bits).
__global__ void max_bits_kernel(const double* __restrict__ vals, size_t n,
                                unsigned long long* __restrict__ out) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    unsigned long long m = 0;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        m = max(m, (unsigned long long)__double_as_longlong(vals[v]));
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned long long)__shfl_xor((long long)m, o, 64));
    if (lane_id() == 0 && m) atomicMax(out, m);
}

// vals[v] *= flag[v] ? f_flagged : f_other (flag == nullptr: f_other).
__global__ void scale_flagged_kernel(double* __restrict__ vals, size_t n,
                                     const uint8_t* __restrict__ flag, double f_flagged,
                                     double f_other) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        vals[v] *= (flag && flag[v]) ? f_flagged : f_other;
}

inline unsigned blocks_for_threads(size_t threads) {
    size_t b = (threads + kBlock - 1) / kBlock;
    return (unsigned)std::max<size_t>(1, std::min<size_t>(b, 1u << 20));
}
inline unsigned blocks_for_waves(size_t waves) { return blocks_for_threads(waves * 64); }

// Blocks of a grid-stride kernel: enough for `waves` waves, but no more
// than the device holds at once, so waves stay resident and loop over
// their work instead of being launched one per item.
std::mutex g_resident_mu;
std::map<std::pair<const void*, int>, unsigned> g_resident;

// (blocks of `block` threads; dynamic LDS per block `lds` bytes)
template <class K>
unsigned resident_blocks(K kernel, size_t waves, int block = kBlock, size_t lds = 0) {
    unsigned cap;
    {
        std::lock_guard<std::mutex> lock(g_resident_mu);
        const auto key = std::make_pair((const void*)kernel, block);
        auto it = g_resident.find(key);
        if (it == g_resident.end()) {
            int dev = 0, cus = 0, per = 0;
            NETBC_HIP(hipGetDevice(&dev));
            NETBC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            NETBC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, lds));
            cap = (unsigned)std::max(1, cus * std::max(per, 1));
            g_resident.emplace(key, cap);
        } else {
            cap = it->second;
        }
    }
    const size_t wpb = (size_t)std::max(block, 64) / 64;
    const size_t b = (waves + wpb - 1) / wpb;
    return (unsigned)std::max<size_t>(1, std::min<size_t>({b, (size_t)1 << 20, (size_t)cap}));
}

// Compute units of the current device.
inline unsigned cu_count() {
    int dev = 0, cus = 0;
    NETBC_HIP(hipGetDevice(&dev));
    NETBC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return (unsigned)std::max(cus, 1);
}

// Grid of a kernel that splits its work by XCD (xcd_wave): a multiple of 8.
inline unsigned xcd_grid(unsigned blocks) {
    return std::max<unsigned>(kXcds, (blocks + kXcds - 1) / kXcds * kXcds);
}

// Grid of the ticketed sweep: a multiple of kTicketClasses, at least one
// workgroup per class.
inline unsigned ticket_grid(unsigned blocks) {
    return std::max<unsigned>(kTicketClasses,
                              (blocks + kTicketClasses - 1) / kTicketClasses * kTicketClasses);
}

// Wavefronts of `kernel` that fit on the device at once.
template <class K>
size_t resident_waves(K kernel, int block = kBlock, size_t lds = 0) {
    return (size_t)resident_blocks(kernel, ~(size_t)0 >> 8, block, lds) * (block / 64);
}

// Dynamic LDS of the kernels that stage queue entries per wavefront
// (QueueStage): kPullBuf 16-byte entries per wavefront of a `block`.
inline size_t pull_lds_bytes(int block) { return (size_t)(block / 64) * kPullBuf * 16; }

// Durations of one kernel's launches from pairs of events recorded around
// each launch on the stream it runs on; read back only once the stream has
// been synchronised, so timing adds no host waits.
struct KernelTimer {
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    double ms = 0.0;
    int64_t launches = 0;
    void mark(hipStream_t st) {
        if (used == pool.size()) {
            hipEvent_t e;
            NETBC_HIP(hipEventCreate(&e));
            pool.push_back(e);
        }
        NETBC_HIP(hipEventRecord(pool[used++], st));
    }
    // Call only after the stream has been synchronised.
    void collect() {
        for (size_t i = 0; i + 1 < used; i += 2) {
            float m = 0.0f;
            NETBC_HIP(hipEventElapsedTime(&m, pool[i], pool[i + 1]));
            ms += m;
            launches++;
        }
        used = 0;
    }
    ~KernelTimer() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

// Device scratch of the batched engine, cached on the graph handle.
struct Workspace {
    int slots = 0;
    uint64_t n = 0;
    DevBuf<uint64_t> vis, cand;
    // Level bitmaps: kLevelBufs buffers of slots * n words. Levels below
    // kKeptLevels keep their own buffer from the forward search to the
    // sweep; deeper levels alternate between the last two buffers and are
    // rebuilt from the level queue during the sweep.
    DevBuf<uint64_t> levels;
    size_t zero_words = 0;  // leading words of `levels` that may hold bits
    DevBuf<uint64_t> succ;  // lanes in which a node has a successor on the next level
    DevBuf<uint64_t> flanes;  // per level parity and slot: lanes with entries on the level
    DevBuf<uint64_t> inb;  // integer lengths: pending-bucket bitmasks per (slot, node)
    DevBuf<double> sig, coef64, acc;
    DevBuf<double> bsum;  // free-order breadth-first runs: dependency sums per (slot, node)
    DevBuf<float> coef32;
    DevBuf<uint32_t> dist;   // integer lengths: distance rows u32[n][64]
    DevBuf<double> rdist;    // real lengths (bucketed): distance rows f64[n][64]
    DevBuf<int64_t> gcur, list, list2;
    DevBuf<QEnt> q;
    DevBuf<uint32_t> hq;    // positions of heavy-row entries in q
    DevBuf<uint32_t> hpre;  // chunk-task prefix over hq (per level)
    DevBuf<uint32_t> perm;  // a level's entry positions in slot order (sweep)
    DevBuf<uint32_t> slot_cnt;  // entries per slot, then their first positions
    DevBuf<double> part;  // per-chunk partial sums of heavy rows
    // Edge runs: per-arc credit, one row of arcs per slot (per thread for
    // real lengths), summed over the rows at the end (fold_arcs_kernel).
    DevBuf<double> eacc;
    DevBuf<uint32_t> npr;  // quirk runs: predecessor arcs per (node, lane), zero between runs
    // Integer lengths on the distance ring: W bitmaps of settled lanes and W
    // of offered lanes, slots * n words each (zero between super-batches
    // unless ring_dirty), and per distance the level's first queue entry,
    // first heavy-list entry and heavy chunk count (lvl_cap distances).
    DevBuf<uint64_t> ring, cring;
    bool ring_dirty = true;
    DevBuf<uint32_t> lvl_q, lvl_h, hc_lvl;
    size_t lvl_cap = 0;
    size_t qcap = 0;
    DevBuf<uint2> touched;
    DevBuf<uint4> dlist;  // bucketed searches: listed (node, slot, out-row bounds)
    DevBuf<Ctrl> ctrl;
    Ctrl* host_ctrl = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    KernelTimer sweep;  // the dependency sweep (dominant kernel)
    // The path-count rows are zero between runs; a run that stops early
    // (an error) leaves them dirty, and the next run clears them first.
    bool rows_dirty = false;
    void clean_rows(hipStream_t st) {
        if (!rows_dirty) return;
        NETBC_HIP(hipMemsetAsync(sig.p, 0, sig.n * sizeof(double), st));
        rows_dirty = false;
    }
    ~Workspace() {
        if (host_ctrl) (void)hipHostFree(host_ctrl);
        for (hipEvent_t e : {ev0, ev1})
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
    }
    return *static_cast<Workspace*>(g.workspace);
}

unsigned long long spread_sum(const unsigned long long (&c)[kSpread]) {
    unsigned long long t = 0;
    for (int i = 0; i < kSpread; i++) t += c[i];
    return t;
}

void read_ctrl(Workspace& ws, hipStream_t st) {
    NETBC_HIP(hipMemcpyAsync(ws.host_ctrl, ws.ctrl.p, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
    NETBC_HIP(hipStreamSynchronize(st));
}

// Real lengths run as a bucketed search when buckets of half the shortest
// arc keep the window (longest arc / width + 2 buckets) within this many
// and every bucket index below 2^32; otherwise (very unequal lengths) as
// the reference's Dijkstra, one source per thread.
constexpr double kRealWindowMax = 4096.0;

struct Buckets {
    int W = 1;           // buckets in the window (tentative distances span < W)
    int words = 1;       // 64-bit words of a pending mask
    double delta = 1.0;  // bucket width
};

bool real_bucketed(const Lengths& L, uint64_t n) {
    if (L.kind != kReal || !(L.rmin > 0.0)) return false;
    const double delta = L.rmin / 2.0;
    const double window = std::floor(L.rmax / delta) + 2.0;
    const double last = (double)std::max<uint64_t>(n, 1) * L.rmax / delta + window;
    return window <= kRealWindowMax && last < 4.0e9;
}

// (width: integer lengths' distances per bucket, at most the shortest arc)
Buckets buckets_for(const Lengths& L, int width = 1) {
    Buckets b;
    if (L.kind == kInteger) {
        // bucket k holds distances [k w, k w + w); offers from it reach at
        // most k w + w - 1 + max, i.e. (w - 1 + max) / w buckets ahead
        const uint64_t w = (uint64_t)std::max(width, 1);
        b.W = (int)((L.max + w - 1) / w + 1);
        b.delta = (double)w;
    } else if (L.kind == kReal) {
        b.delta = L.rmin / 2.0;
        b.W = (int)(std::floor(L.rmax / b.delta) + 2.0);
    }
    b.words = (b.W + 63) / 64;
    return b;
}

// Integer lengths run on the distance ring (dial_gather_kernel /
// dial_expand_kernel) when the ring is short (longest arc below kRingMaxW)
// and a wavefront's concatenated in-rows stay within 32-bit indices;
// otherwise, and for distance-only searches, on the per-lane distance rows
// (dial_step_kernel).
constexpr uint32_t kRingMaxW = 64;
// Ring-search steps between checks for its end (one host wait each).
constexpr uint32_t kRingCheck = 8;
// Switched off: a GPU run of the random-multigraph parity tests that
// reached this path ended in a device memory fault whose cause is not
// established (a CPU execution of these kernels on the same inputs is in
// bounds and exact), so integer lengths keep the per-lane distance rows.
constexpr bool kRingEnabled = false;
bool ring_path(const Graph& g, const Lengths& L) {
    return kRingEnabled && L.kind == kInteger && L.max + 1 <= kRingMaxW &&
           g.dev.max_in <= (1ull << 25);
}

// Sizes (in bytes per node and slot) of the per-slot arrays.
// bitmaps (visited, offers, successors, the level buffers), the touched
// list, about four queue entries with their heavy-row slots, and the
// per-lane rows.
size_t slot_bytes_per_node(const Lengths& L, int precision, uint64_t n, bool ring = false) {
    // visited/candidate/successor words, touched list, 4 queue entries per
    // pair (entry, heavy-row position, chunk prefix, slot-order position),
    // path-count row, coefficient row
    size_t b = 3 * 8 + 8 + 4 * (sizeof(QEnt) + 12) + kLanes * 8 + kLanes * (precision ? 4 : 8);
    if (L.kind == kUniform) b += kLevelBufs * 8 + 8;  // level bitmaps, dependency sums (bsum)
    if (L.kind == kInteger && ring) b += 2 * (size_t)(L.max + 1) * 8 + 8;  // rings, bsum
    if (L.kind == kInteger && !ring) b += kLanes * 4 + 8 * buckets_for(L).words + 8;
    if (L.kind == kReal) {
        if (real_bucketed(L, n)) b += kLanes * 8 + 8 * buckets_for(L).words + 8;
        else b += kLanes * (4 * 8 + 4 * 4);  // per-thread Dijkstra arrays
    }
    return b;
}

// The tuning table's entry for this graph and length class (tuning.h), or
// nullptr (real lengths, or no class matches).
const TuneEntry* tuning_for(const Graph& g, const Lengths& L) {
    if (L.kind == kReal) return nullptr;
    const int kind = L.kind == kUniform ? 0 : 1;
    int lg = 0;
    while (lg < 63 && (2ull << lg) <= g.n) lg++;
    const double deg = g.n ? (double)g.dev.arcs / (double)g.n : 0.0;
    for (const TuneEntry& e : kTuning) {
        if (e.kind != kind) continue;
        if (e.directed >= 0 && e.directed != (g.directed ? 1 : 0)) continue;
        if (lg < e.log2n_min || lg > e.log2n_max) continue;
        if (deg < e.deg_min || deg >= e.deg_max) continue;
        return &e;
    }
    return nullptr;
}

// Options with the tuned defaults filled in where the caller left 0.
RunOptions tuned_options(const Graph& g, const Lengths& L, const RunOptions& opt) {
    RunOptions o = opt;
    const TuneEntry* t = tuning_for(g, L);
    if (o.slots <= 0)
        o.slots = t ? t->slots : (L.kind == kReal && !real_bucketed(L, g.n) ? 4 : 16);
    if (!(o.alpha > 0.0)) o.alpha = t ? t->alpha : kBottomUpAlpha;
    if (o.sweep_group <= 0) o.sweep_group = t ? t->sweep_group : (int)kSweepGroup;
    if (o.sweep_block <= 0) o.sweep_block = t ? t->sweep_block : kBlock;
    if (o.step_block <= 0) o.step_block = t ? t->step_block : kBlock;
    if (o.hub_in <= 0) o.hub_in = t ? t->hub_in : (int)kHubIn;
    if (o.hub_chunk <= 0) o.hub_chunk = t ? t->hub_chunk : (int)kHubChunk;
    if (o.bucket_width <= 0) o.bucket_width = t ? t->bucket_width : 1;
    // buckets no wider than the shortest arc are settled whole
    if (L.kind == kInteger) o.bucket_width = (int)std::min<uint32_t>(o.bucket_width, L.min);
    else o.bucket_width = 1;
    return o;
}

// Slots per super-batch: the tuned (or requested) count, capped by the
// number of batches and by a 128 GiB share of the device's 288 GB (edge
// runs add a row of per-arc credit per slot, per thread for real lengths),
// then evened out over the super-batches that cap implies, so the last one
// is not a small remainder (a super-batch's levels cost latency whatever
// its width: R-MAT 22 with 8,192 sources, 18 slots as 8 super-batches
// 1.66 s, 16 slots as 8 1.55 s, 22 slots as 6 1.52 s).
int choose_slots(const Graph& g, const Lengths& L, int64_t k, const RunOptions& opt,
                 bool edge = false, bool ring = false) {
    const int64_t batches = (k + kLanes - 1) / kLanes;
    if (batches <= 0) return 1;
    int want = std::min(std::max(opt.slots, 1), 1024);
    size_t per = slot_bytes_per_node(L, opt.precision, g.n, ring) * std::max<uint64_t>(g.n, 1);
    const bool per_thread = L.kind == kReal && !real_bucketed(L, g.n);
    if (edge) per += (size_t)g.dev.arcs * sizeof(double) * (per_thread ? kLanes : 1);
    size_t budget = (size_t)128 << 30;
    int by_mem = (int)std::max<size_t>(1, budget / std::max<size_t>(per, 1));
    int s = std::max<int>(1, (int)std::min<int64_t>({(int64_t)want, batches, (int64_t)by_mem}));
    const int64_t supers = (batches + s - 1) / s;
    return (int)std::max<int64_t>(1, (batches + supers - 1) / supers);
}

}  // namespace

void run_batched(Graph& g, const Lengths& L, const int64_t* sources, int64_t k,
                 const RunOptions& opt, double* d_bc, RunDiag& diag, double* d_eb,
                 const SubsetRun* sub);

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
    bool edge = false;  // add every arc's credit to ws.eacc
    const uint8_t* target = nullptr;  // target flags (nullptr: every node)
    bool tie_exact = false;           // real lengths tie only when equal
    bool quirk = false;               // NetworkX's subset edge split (edge runs)
    uint32_t* npr = nullptr;          // predecessor-arc rows (quirk runs)
    bool ring = false;                // integer lengths on the distance ring
    uint32_t ring_top = 0;            // last distance the ring search stepped
    // Free-order breadth-first runs: per (slot, node) sums of the swept
    // dependencies (ws.bsum), folded into the scores after each super-batch.
    double* bsum = nullptr;
    // Distance-profile runs (breadth-first): per-lane profiles of the
    // super-batch (5 * slots * kLanes words from lane_buf, cleared at the
    // start of every search) and per-lane level counts.
    uint64_t* lane_buf = nullptr;
    unsigned long long* lane_count = nullptr;
    bool reach_only = false;  // distances only, no path counts
    // Reachability-only breadth-first runs time their bottom-up steps with
    // ws.sweep; bytes of the stale level entries those steps clear (24 each).
    unsigned long long stale_bytes = 0;

    // Per-level boundaries in the queue (level d = [lv[d], lv[d+1])).
    std::vector<uint32_t> lv;
    std::vector<uint32_t> hv;       // heavy-row list boundaries per level
    std::vector<uint32_t> hc;       // chunk tasks of each level's heavy rows

    // Prefix of chunk counts for the heavy rows listed from hbeg on.
    void scan_heavy(uint32_t hbeg) {
        heavy_scan_kernel<<<1, 1024, 0, st>>>(ws.hq.p, hbeg, &ws.ctrl.p->hlen, ws.q.p,
                                              g.dev.out_off.p, ws.hpre.p, &ws.ctrl.p->hchunks);
        launch_check();
    }

    void launch_check() { NETBC_HIP(hipGetLastError()); }

    void clear_rows() {
        clear_rows_kernel<<<blocks_for_threads((size_t)slots * n), kBlock, 0, st>>>(
            ws.vis.p, ws.sig.p, (size_t)slots * n);
        launch_check();
    }

    void reset_slots() {
        size_t nm = (size_t)slots * n;
        unsigned gb = blocks_for_threads(std::min<size_t>(nm, (size_t)1 << 22));
        if (L.kind == kUniform) {
            zero3_kernel<<<blocks_for_threads(std::min<size_t>(nm / 2 + 1, (size_t)1 << 22)),
                           kBlock, 0, st>>>(ws.vis.p, ws.cand.p, ws.succ.p, nm);
            launch_check();
            if (ws.zero_words > 0)
                fill_u64_kernel<<<blocks_for_threads(std::min<size_t>(ws.zero_words,
                                                                      (size_t)1 << 22)),
                                  kBlock, 0, st>>>(ws.levels.p, ws.zero_words, 0);
            ws.zero_words = 0;
            launch_check();
        } else {
            fill_u64_kernel<<<gb, kBlock, 0, st>>>(ws.vis.p, nm, 0);
            launch_check();
        }
        const unsigned rb = blocks_for_threads(std::min<size_t>(nm * kLanes, (size_t)1 << 22));
        if (L.kind == kInteger)
            fill_u32_kernel<<<rb, kBlock, 0, st>>>(ws.dist.p, nm * kLanes, kNone);
        else if (L.kind == kReal && real_bucketed(L, n))
            fill_u64_kernel<<<rb, kBlock, 0, st>>>((uint64_t*)ws.rdist.p, nm * kLanes,
                                                   RealPath::kInfBits);
        if (L.kind == kInteger || (L.kind == kReal && real_bucketed(L, n)))
            fill_u64_kernel<<<gb, kBlock, 0, st>>>(ws.inb.p, nm * buckets_for(L).words, 0);
        launch_check();
        NETBC_HIP(hipMemsetAsync(ws.ctrl.p, 0, offsetof(Ctrl, sig_overflow), st));
    }

    template <class C>
    C* coef_ptr();

    bool run_bfs(int64_t item0) {
        const DeviceGraph& d = g.dev;
        // a super-batch repeated after a queue overflow counts its levels once
        const int64_t levels0 = diag.levels;
        const int32_t bottom_up0 = diag.bottom_up_levels;
        reset_slots();
        if (lane_buf) {
            const size_t words = (size_t)5 * slots * kLanes;
            fill_u64_kernel<<<blocks_for_threads(words), kBlock, 0, st>>>(lane_buf, words, 0);
            launch_check();
        }
        init_sources_kernel<uint32_t><<<blocks_for_threads((size_t)slots * kLanes), kBlock, 0, st>>>(
            d_list, item0, k, slots, n, ws.vis.p, level_bits(0), ws.sig.p, nullptr, nullptr, 0);
        launch_check();
        source_entries_kernel<<<blocks_for_threads((size_t)slots * kLanes), kBlock, 0, st>>>(
            d_list, item0, k, slots, n, ws.vis.p, ws.q.p, ws.ctrl.p, d.heavy_flag.p, ws.hq.p,
            d.out_off.p);
        launch_check();
        read_ctrl(ws, st);
        ws.sweep.collect();
        lv.assign(1, 0);
        lv.push_back(ws.host_ctrl->qlen);
        hv.assign(1, 0);
        hc.clear();
        scan_heavy(0);
        frontier(0, 0, lv[1]);
        read_ctrl(ws, st);
        hv.push_back(ws.host_ctrl->hlen);
        hc.push_back(ws.host_ctrl->hchunks);
        const int64_t left = k - item0;  // sources from lane 0 of slot 0 on
        const uint64_t live_slots =
            (uint64_t)std::min<int64_t>(slots, (left + kLanes - 1) / kLanes);
        const size_t pairs = (size_t)slots * n;
        const double alpha = opt.alpha > 0.0 ? opt.alpha : kBottomUpAlpha;
        unsigned long long* front[2] = {ws.ctrl.p->front_arcs[0], ws.ctrl.p->front_arcs[1]};
        for (int dl = 0;; dl++) {
            uint32_t qb = lv[dl], qe = lv[dl + 1];
            uint32_t hb = hv[dl], he = hv[dl + 1];
            // The next level's buffer holds level dl-1 only where buffers
            // alternate; those stale entries are cleared during the step.
            const bool stale = dl - 1 >= kKeptLevels;
            uint32_t zb = stale ? lv[dl - 1] : 0, ze = stale ? lv[dl] : 0;
            uint64_t* cur = level_bits(dl);
            uint64_t* nxt = level_bits(dl + 1);
            // Direction of this step: top-down costs the out-arcs of the
            // level's entries, bottom-up the in-arcs of the pairs with
            // unreached active lanes.
            const double f_arcs = (double)spread_sum(ws.host_ctrl->front_arcs[dl & 1]);
            const double open_arcs =
                (double)live_slots * (double)d.arcs - (double)spread_sum(ws.host_ctrl->closed_arcs);
            const bool bottom_up = opt.direction == 2 ||
                                   (opt.direction == 0 && f_arcs * alpha > open_arcs);
            if (bottom_up) {
                diag.bottom_up_levels++;
                auto pull = reach_only ? bfs_pull_kernel<false>
                            : npr      ? bfs_pull_kernel<true, true>
                                       : bfs_pull_kernel<true, false>;
                if (reach_only) {
                    ws.sweep.mark(st);
                    stale_bytes += 24ull * (ze - zb);
                }
                const int pb = opt.step_block;
                pull<<<xcd_grid(resident_blocks(pull, (pairs + 63) / 64, pb, pull_lds_bytes(pb))),
                       pb, pull_lds_bytes(pb), st>>>(ws.q.p, zb, ze, d.in_off_p(), d.in_idx_p(), d.out_off.p,
                                        ws.vis.p, cur, nxt, ws.succ.p, ws.sig.p,
                                        d.hub_flag.p, d.heavy_flag.p, ws.q.p, (uint32_t)ws.qcap,
                                        ws.hq.p, ws.ctrl.p, front[(dl + 1) & 1], front[dl & 1],
                                        frontier_lanes(dl), slots, left, n, npr);
                launch_check();
                if (reach_only) ws.sweep.mark(st);
            } else {
                const uint32_t gsz =
                    group_size(qe - qb, resident_waves(bfs_expand_kernel));
                size_t tasks = (size_t)(qe - qb + gsz - 1) / gsz + hc[dl];
                size_t waves = std::max<size_t>(tasks, (ze - zb + 63) / 64);
                bfs_expand_kernel<<<ticket_grid(resident_blocks(bfs_expand_kernel, waves)), kBlock,
                                    0, st>>>(
                    ws.q.p, qb, qe, zb, ze, ws.hq.p, hb, he, ws.hpre.p, hc[dl], d.out_off.p,
                    d.out_idx.p, ws.vis.p, ws.cand.p, nxt, reach_only ? nullptr : ws.succ.p,
                    ws.touched.p,
                    &ws.ctrl.p->tlen[dl & 1], d.heavy_flag.p, ws.ctrl.p, front[dl & 1], gsz, n);
                launch_check();
            }
            if (d.hubs && reach_only && bottom_up) {
                const size_t hub_tasks = (size_t)d.hub_chunks * slots;
                reach_hub_kernel<<<resident_blocks(reach_hub_kernel, hub_tasks), kBlock, 0, st>>>(
                    d.hub_task.p, d.hub_chunks, slots, d.in_idx_p(), cur, ws.cand.p, ws.vis.p,
                    ws.touched.p, &ws.ctrl.p->tlen[dl & 1], frontier_lanes(dl), n);
                launch_check();
            } else if (d.hubs && !reach_only) {
                const size_t hub_tasks = (size_t)d.hub_chunks * slots;
                auto hub = npr ? bfs_hub_kernel<true> : bfs_hub_kernel<false>;
                const int hb = opt.step_block;
                const uint32_t gsz = group_size(hub_tasks, resident_waves(hub, hb));
                hub<<<xcd_grid(resident_blocks(hub, (hub_tasks + gsz - 1) / gsz, hb)), hb, 0,
                      st>>>(d.hub_task.p, d.hub_chunks, slots, d.in_idx_p(),
                                                  cur, ws.cand.p, ws.sig.p, bottom_up ? 1 : 0,
                                                  ws.vis.p, ws.succ.p, ws.touched.p,
                                                  &ws.ctrl.p->tlen[dl & 1],
                                                  frontier_lanes(dl), ws.ctrl.p, gsz, n, npr);
                launch_check();
            }
            auto gather = reach_only ? bfs_gather_kernel<false>
                          : npr      ? bfs_gather_kernel<true, true>
                                     : bfs_gather_kernel<true, false>;
            gather<<<ticket_grid(resident_blocks(gather, (pairs + 63) / 64)), kBlock, kPullLds,
                     st>>>(
                ws.touched.p, &ws.ctrl.p->tlen[dl & 1],
                                      &ws.ctrl.p->tlen[(dl + 1) & 1], d.in_off_p(), d.in_idx_p(),
                                      d.out_off.p, cur, ws.cand.p, ws.vis.p, nxt,
                                      ws.sig.p, d.hub_flag.p, d.heavy_flag.p, ws.q.p,
                                      (uint32_t)ws.qcap, ws.hq.p, ws.ctrl.p, left,
                                      front[(dl + 1) & 1], n, npr);
            launch_check();
            scan_heavy(hv.back());
            read_ctrl(ws, st);
            if (ws.host_ctrl->qoverflow) {
                ws.zero_words = ws.levels.n;
                diag.levels = levels0;
                diag.bottom_up_levels = bottom_up0;
                return false;
            }
            uint32_t ql = ws.host_ctrl->qlen;
            if (ql == qe) break;
            if (lane_buf) count_level(nxt, (uint32_t)dl + 1);
            lv.push_back(ql);
            hv.push_back(ws.host_ctrl->hlen);
            hc.push_back(ws.host_ctrl->hchunks);
            frontier(dl + 1, qe, ql);
            diag.levels++;
        }
        // buffers written: those of levels 0 .. D+1 (D: last level with entries)
        const int D = (int)lv.size() - 2;
        ws.zero_words = (size_t)std::min(D + 2, kLevelBufs) * slots * n;
        return true;
    }

    // Adds breadth-first level `level` (bitmap `bits`) to the lane profiles.
    void count_level(const uint64_t* bits, uint32_t level) {
        endpoint_lanes_kernel<<<resident_blocks(endpoint_lanes_kernel,
                                                (n + 63) / 64 * (size_t)slots),
                                kBlock, 0, st>>>(bits, slots, n, nullptr, lane_count);
        launch_check();
        level_profile_kernel<<<blocks_for_threads((size_t)slots * kLanes), kBlock, 0, st>>>(
            lane_count, slots, level, profile_view(lane_buf, (size_t)slots * kLanes));
        launch_check();
    }

    // Lane profiles of a finished bucketed search, from its distance rows.
    template <class P>
    void dist_profile(DevProfile lanes, uint32_t* pc, unsigned long long* pit, double* pft,
                      double* ph, double* pl) {
        const size_t chunks = (n + kProfileChunk - 1) / kProfileChunk;
        dist_profile_kernel<P><<<resident_blocks(dist_profile_kernel<P>, chunks * slots), kBlock,
                                 0, st>>>(dist_rows<P>(), slots, n, pc, pit, pft, ph, pl);
        launch_check();
        dist_profile_reduce_kernel<<<blocks_for_threads((size_t)slots * kLanes), kBlock, 0, st>>>(
            pc, pit, pft, ph, pl, slots, n, lanes);
        launch_check();
    }

    // Lanes with entries on level d, per slot (flanes, by level parity).
    uint64_t* frontier_lanes(int d) { return ws.flanes.p + (size_t)(d & 1) * slots; }
    void frontier(int d, uint32_t qb, uint32_t qe) {
        if (opt.direction == 1) return;  // top-down only: not needed
        uint64_t* fl = frontier_lanes(d);
        NETBC_HIP(hipMemsetAsync(fl, 0, (size_t)slots * sizeof(uint64_t), st));
        if (qe > qb) {
            unsigned blocks = std::min<unsigned>(256, blocks_for_threads(qe - qb));
            frontier_lanes_kernel<<<blocks, kBlock, (size_t)slots * sizeof(uint64_t), st>>>(
                ws.q.p, qb, qe, fl, slots);
            launch_check();
        }
    }

    // Bitmap buffer of level d (see Workspace::levels).
    uint64_t* level_bits(int d) {
        const int b = d < kKeptLevels ? d : kKeptLevels + ((d - kKeptLevels) & 1);
        return ws.levels.p + (size_t)b * slots * n;
    }

    // The backward sweep of the super-batch just searched.
    template <class C>
    void backward(C* coef) {
        if (ring) {
            if (quirk) backward_ring<C, true, true>(coef);
            else if (edge) backward_ring<C, true>(coef);
            else backward_ring<C, false>(coef);
            return;
        }
        if (quirk) {
            if (L.kind == kUniform) backward_bfs<C, true, true>(coef);
            else if (L.kind == kInteger) backward_dial<C, true, IntPath, true>(coef);
            else if (tie_exact) backward_dial<C, true, RealPathExact, true>(coef);
            else backward_dial<C, true, RealPath, true>(coef);
        } else if (L.kind == kUniform) {
            if (edge) backward_bfs<C, true>(coef);
            else backward_bfs<C, false>(coef);
        } else if (L.kind == kInteger) {
            if (edge) backward_dial<C, true, IntPath>(coef);
            else backward_dial<C, false, IntPath>(coef);
        } else if (tie_exact) {
            if (edge) backward_dial<C, true, RealPathExact>(coef);
            else backward_dial<C, false, RealPathExact>(coef);
        } else {
            if (edge) backward_dial<C, true, RealPath>(coef);
            else backward_dial<C, false, RealPath>(coef);
        }
    }

    // Distance rows and lengths of a bucketed search.
    template <class P>
    typename P::Dist* dist_rows() {
        if constexpr (std::is_same_v<P, IntPath>) return ws.dist.p;
        else return ws.rdist.p;
    }
    template <class P>
    const typename P::Len* out_lengths() {
        if constexpr (std::is_same_v<P, IntPath>) return L.out_len.p;
        else return L.out_real.p;
    }
    template <class P>
    const typename P::Len* in_lengths() {
        if constexpr (std::is_same_v<P, IntPath>) return g.dev.directed ? L.in_len.p : L.out_len.p;
        else return g.dev.directed ? L.in_real.p : L.out_real.p;
    }

    // perm[qb, qe) = the level's entry positions grouped by slot
    // (slot_count / slot_scan / slot_scatter kernels).
    void order_by_slot(uint32_t qb, uint32_t qe) {
        const size_t cnt = qe - qb;
        const unsigned blocks = blocks_for_threads(std::min<size_t>(cnt, (size_t)1 << 18));
        fill_u32_kernel<<<1, 256, 0, st>>>(ws.slot_cnt.p, (size_t)slots, 0);
        slot_count_kernel<<<blocks, kBlock, (size_t)slots * 4, st>>>(ws.q.p, qb, qe, slots,
                                                                    ws.slot_cnt.p);
        slot_scan_kernel<<<1, 1024, 0, st>>>(ws.slot_cnt.p, slots);
        slot_scatter_kernel<<<blocks, kBlock, (size_t)slots * 8, st>>>(ws.q.p, qb, qe, slots,
                                                                      ws.slot_cnt.p, ws.perm.p);
        launch_check();
    }

    template <class C, bool kEdge, bool kQuirk = false>
    void backward_bfs(C* coef) {
        const DeviceGraph& d = g.dev;
        int D = (int)lv.size() - 2;  // last level with entries
        const int split = opt.exact || kQuirk ? 0 : 1;
        // edge runs also sweep the sources: their out-arcs carry credit
        for (int dl = D; dl >= (kEdge ? 0 : 1); dl--) {
            uint32_t qb = lv[dl], qe = lv[dl + 1];
            uint32_t hb = hv[dl], he = hv[dl + 1];
            // Levels in alternating buffers are rebuilt for the next step:
            // W = level dl, from the buffer holding level dl+2.
            const bool rebuild = dl >= kKeptLevels && dl < D;
            uint32_t ab = rebuild && dl + 2 <= D ? lv[dl + 2] : 0;
            uint32_t ae = rebuild && dl + 2 <= D ? lv[dl + 3] : 0;
            bool heavy = split && dl < D && he > hb;
            const int sb = opt.sweep_block;
            uint32_t gsz =
                group_size(qe - qb, resident_waves(bfs_backward_kernel<C, kEdge, kQuirk>, sb),
                           (uint32_t)std::min(std::max(opt.sweep_group, 1), 64));
            // A task walks its entries' rows as one concatenation indexed in
            // 32 bits: with whole rows (no split) of up to 2^32 - 2 arcs,
            // one entry per task keeps it in range.
            if (!split && d.max_out * gsz > 0xFFFFFFFFull - 64) gsz = 1;
            size_t tasks = (size_t)(qe - qb + gsz - 1) / gsz + (heavy ? hc[dl] : 0);
            size_t waves = std::max<size_t>(tasks, rebuild ? (qe - qb + ae - ab + 63) / 64 : 0);
            if (heavy && ws.part.n < (size_t)hc[dl] * kLanes) {
                NETBC_HIP(hipStreamSynchronize(st));  // the old buffer may be in use
                ws.part.alloc((size_t)hc[dl] * kLanes * 2);
            }
            // large levels are swept slot by slot (entries with successors)
            const uint32_t* perm = nullptr;
            if (dl < D && slots > 1 && qe - qb >= kSlotOrderMin) {
                order_by_slot(qb, qe);
                perm = ws.perm.p;
            }
            ws.sweep.mark(st);
            bfs_backward_kernel<C, kEdge, kQuirk>
                <<<ticket_grid(resident_blocks(bfs_backward_kernel<C, kEdge, kQuirk>, waves, sb)),
                   sb, 0, st>>>(
                    ws.q.p, qb, qe, ab, ae, dl == D, rebuild, ws.hq.p, hb, he, ws.hpre.p, hc[dl],
                    split, d.heavy_flag.p, ws.part.p, d.out_off.p, d.out_idx.p,
                    LevelBits{level_bits(dl + 1)}, level_bits(dl), ws.succ.p, ws.sig.p, coef, ws.ctrl.p, gsz,
                    n, ws.eacc.p, d.arcs, target, npr, dl > 0 ? bsum : nullptr, perm);
            launch_check();
            if (heavy) {
                bfs_combine_kernel<C><<<resident_blocks(bfs_combine_kernel<C>, he - hb), kBlock, 0, st>>>(
                    ws.q.p, ws.hq.p, hb, he, ws.hpre.p, hc[dl], ws.part.p, d.out_off.p, ws.sig.p,
                    coef, ws.ctrl.p, n, target, dl > 0 ? bsum : nullptr);
                launch_check();
            }
            ws.sweep.mark(st);
        }
    }

    // Bucketed search (Dial for integer lengths, P = IntPath; buckets of
    // half the shortest arc for real lengths, P = RealPath): one scan and
    // one step per bucket, one host sync per bucket.
    template <class P>
    bool run_bucketed(int64_t item0) {
        const DeviceGraph& d = g.dev;
        const int64_t levels0 = diag.levels;  // counted once if the super-batch repeats
        reset_slots();
        const Buckets bk = buckets_for(L, opt.bucket_width);
        const int W = bk.W;
        const int words = bk.words;
        typename P::Dist* dist = dist_rows<P>();
        init_sources_kernel<<<blocks_for_threads((size_t)slots * kLanes), kBlock, 0, st>>>(
            d_list, item0, k, slots, n, ws.vis.p, nullptr, ws.sig.p, dist, ws.inb.p, words);
        launch_check();
        source_entries_kernel<<<blocks_for_threads((size_t)slots * kLanes), kBlock, 0, st>>>(
            d_list, item0, k, slots, n, ws.vis.p, ws.q.p, ws.ctrl.p, d.heavy_flag.p, nullptr,
            nullptr);
        launch_check();
        read_ctrl(ws, st);
        ws.sweep.collect();
        lv.assign(1, 0);
        lv.push_back(ws.host_ctrl->qlen);
        const size_t pairs = (size_t)slots * n;
        const unsigned scan_blocks = std::min(
            resident_blocks(dial_scan_kernel, (pairs + 64 * kScanUnroll - 1) / (64 * kScanUnroll)),
            cu_count());
        auto step = d.directed ? (reach_only ? dial_step_kernel<P, false, false>
                                             : dial_step_kernel<P, true, false>)
                               : (reach_only ? dial_step_kernel<P, false, true>
                                             : dial_step_kernel<P, true, true>);
        const unsigned step_blocks = resident_blocks(step, pairs);
        // the first step's bucket minimum starts empty; later ones are reset
        // by the scan before them
        NETBC_HIP(hipMemsetAsync(&ws.ctrl.p->next_bucket[0], 0xFF, sizeof(unsigned long long),
                                 st));
        uint32_t D = 0;
        for (uint32_t it = 0;; it++) {
            const int p = it & 1;
            dial_scan_kernel<<<scan_blocks, kBlock, 0, st>>>(
                ws.inb.p, words, W, D, pairs, d.out_off.p, ws.dlist.p, &ws.ctrl.p->tlen[p],
                &ws.ctrl.p->tlen[p ^ 1], &ws.ctrl.p->tlong[p], &ws.ctrl.p->tlong[p ^ 1],
                &ws.ctrl.p->left[p], &ws.ctrl.p->left[p ^ 1],
                &ws.ctrl.p->next_bucket[p], &ws.ctrl.p->next_bucket[p ^ 1], n);
            launch_check();
            step<<<ticket_grid(step_blocks), kBlock, 0, st>>>(
                ws.dlist.p, &ws.ctrl.p->tlen[p], &ws.ctrl.p->tlong[p], pairs, D, W, words, D == 0, bk.delta, ws.inb.p,
                d.in_off_p(), d.in_idx_p(), in_lengths<P>(), d.out_off.p, d.out_idx.p,
                out_lengths<P>(), dist, ws.vis.p, ws.sig.p, ws.q.p, (uint32_t)ws.qcap, ws.ctrl.p,
                &ws.ctrl.p->next_bucket[p], n, npr);
            launch_check();
            read_ctrl(ws, st);
            if (ws.host_ctrl->qoverflow) {
                diag.levels = levels0;
                return false;
            }
            const uint32_t ql = ws.host_ctrl->qlen;
            if (ql > lv.back()) {
                lv.push_back(ql);
                diag.levels++;
            }
            // Done when this step listed nothing and no bucket is pending.
            if (ws.host_ctrl->tlen[p] == 0 && ws.host_ctrl->tlong[p] == 0 &&
                ws.host_ctrl->left[p] == 0)
                break;
            // otherwise on to the lowest pending bucket (empty ones skipped)
            const unsigned long long nb = ws.host_ctrl->next_bucket[p];
            if (nb > 0xFFFFFFFFull && nb != ~0ull) throw Error("bucket index exceeds 2^32");
            D = nb != ~0ull && nb > D ? (uint32_t)nb : D + 1;
        }
        return true;
    }

    template <class C, bool kEdge, class P, bool kQuirk = false>
    void backward_dial(C* coef) {
        const DeviceGraph& d = g.dev;
        int D = (int)lv.size() - 2;
        for (int dl = D; dl >= (kEdge ? 0 : 1); dl--) {
            uint32_t qb = lv[dl], qe = lv[dl + 1];
            ws.sweep.mark(st);
            dial_backward_kernel<C, kEdge, P, kQuirk>
                <<<ticket_grid(resident_blocks(dial_backward_kernel<C, kEdge, P, kQuirk>,
                                               qe - qb)),
                   kBlock, 0, st>>>(ws.q.p, qb, qe, d.out_off.p, d.out_idx.p, out_lengths<P>(),
                                    dist_rows<P>(), ws.vis.p, ws.sig.p, coef, n, ws.eacc.p,
                                    d.arcs, target, npr, ws.ctrl.p);
            launch_check();
            ws.sweep.mark(st);
        }
    }

    // Distance tables of the ring search grown to `need` distances (the
    // stream is drained first: the old tables are freed).
    void grow_levels(size_t need) {
        if (need <= ws.lvl_cap) return;
        size_t cap = ws.lvl_cap;
        while (cap < need) cap *= 2;
        for (DevBuf<uint32_t>* b : {&ws.lvl_q, &ws.lvl_h, &ws.hc_lvl}) {
            DevBuf<uint32_t> t;
            t.alloc(cap);
            NETBC_HIP(hipMemcpyAsync(t.p, b->p, ws.lvl_cap * sizeof(uint32_t),
                                     hipMemcpyDeviceToDevice, st));
            NETBC_HIP(hipStreamSynchronize(st));
            std::swap(b->p, t.p);
            std::swap(b->n, t.n);
        }
        ws.lvl_cap = cap;
    }

    // Integer lengths on the distance ring (dial_gather_kernel): the
    // forward search of one super-batch, one step per distance, chained
    // without host waits and checked for its end every kRingCheck steps.
    // Returns false if the level queue overflowed.
    bool run_ring(int64_t item0) {
        const DeviceGraph& d = g.dev;
        const uint32_t W = L.max + 1;
        const size_t stride = (size_t)slots * n;
        const unsigned fb = blocks_for_threads(std::min<size_t>(stride, (size_t)1 << 22));
        fill_u64_kernel<<<fb, kBlock, 0, st>>>(ws.vis.p, stride, 0);
        fill_u64_kernel<<<fb, kBlock, 0, st>>>(ws.succ.p, stride, 0);
        if (ws.ring_dirty) {
            const unsigned rb = blocks_for_threads(std::min<size_t>(stride * W, (size_t)1 << 22));
            fill_u64_kernel<<<rb, kBlock, 0, st>>>(ws.ring.p, stride * W, 0);
            fill_u64_kernel<<<rb, kBlock, 0, st>>>(ws.cring.p, stride * W, 0);
            ws.ring_dirty = false;
        }
        launch_check();
        NETBC_HIP(hipMemsetAsync(ws.ctrl.p, 0, offsetof(Ctrl, sig_overflow), st));
        init_sources_kernel<uint32_t><<<blocks_for_threads((size_t)slots * kLanes), kBlock, 0, st>>>(
            d_list, item0, k, slots, n, ws.vis.p, ws.ring.p, ws.sig.p, nullptr, nullptr, 0);
        launch_check();
        source_entries_kernel<<<blocks_for_threads((size_t)slots * kLanes), kBlock, 0, st>>>(
            d_list, item0, k, slots, n, ws.vis.p, ws.q.p, ws.ctrl.p, d.heavy_flag.p, ws.hq.p,
            nullptr);
        launch_check();
        fill_u32_kernel<<<1, 64, 0, st>>>(ws.lvl_q.p, 1, 0);
        fill_u32_kernel<<<1, 64, 0, st>>>(ws.lvl_h.p, 1, 0);
        launch_check();
        auto gather = npr ? dial_gather_kernel<true> : dial_gather_kernel<false>;
        const unsigned gb = xcd_grid(resident_blocks(gather, (stride + 63) / 64));
        const unsigned eb = resident_blocks(dial_expand_kernel, ~(size_t)0 >> 8);
        uint32_t D = 0;
        for (;; D++) {
            if ((size_t)D + 2 > ws.lvl_cap) grow_levels((size_t)D + 2);
            if (D > 0) {
                gather<<<gb, kBlock, kPullLds, st>>>(
                    D, W, ws.cring.p, ws.ring.p, stride, d.in_off_p(), d.in_idx_p(),
                    in_lengths<IntPath>(), ws.vis.p, ws.succ.p, ws.sig.p, d.heavy_flag.p, ws.q.p,
                    (uint32_t)ws.qcap, ws.hq.p, ws.ctrl.p, slots, n, npr);
                launch_check();
            }
            heavy_scan_kernel<<<1, 1024, 0, st>>>(ws.hq.p, 0, &ws.ctrl.p->hlen, ws.q.p,
                                                  d.out_off.p, ws.hpre.p, &ws.ctrl.p->hchunks,
                                                  ws.lvl_h.p + D, ws.hc_lvl.p + D);
            launch_check();
            dial_expand_kernel<<<eb, kBlock, 0, st>>>(
                D, W, ws.q.p, ws.lvl_q.p, ws.lvl_h.p, ws.hq.p, ws.hpre.p, d.out_off.p,
                d.out_idx.p, out_lengths<IntPath>(), ws.vis.p, ws.cring.p, ws.ring.p, stride,
                d.heavy_flag.p, ws.ctrl.p, n, (uint32_t)ws.qcap);
            launch_check();
            if (D % kRingCheck == kRingCheck - 1) {
                read_ctrl(ws, st);
                if (ws.host_ctrl->qoverflow) {
                    ws.ring_dirty = true;
                    return false;
                }
                if (ws.host_ctrl->last_offer <= D) break;  // nothing offered beyond D
            }
        }
        std::vector<uint32_t> bq((size_t)D + 2), bh((size_t)D + 2), bc((size_t)D + 1);
        NETBC_HIP(hipMemcpyAsync(bq.data(), ws.lvl_q.p, bq.size() * 4, hipMemcpyDeviceToHost, st));
        NETBC_HIP(hipMemcpyAsync(bh.data(), ws.lvl_h.p, bh.size() * 4, hipMemcpyDeviceToHost, st));
        NETBC_HIP(hipMemcpyAsync(bc.data(), ws.hc_lvl.p, bc.size() * 4, hipMemcpyDeviceToHost,
                                 st));
        NETBC_HIP(hipStreamSynchronize(st));
        uint32_t last = 0;  // last distance with entries
        for (uint32_t x = 0; x <= D; x++)
            if (bq[x + 1] > bq[x]) last = x;
        lv.assign(bq.begin(), bq.begin() + last + 2);
        hv.assign(bh.begin(), bh.begin() + last + 2);
        hc.assign(bc.begin(), bc.begin() + last + 1);
        ring_top = D;
        diag.levels += last;
        return true;
    }

    // Clears the ring words of the entries of distances [from, to] (each
    // ring slot holds one distance, so a word is that distance's lanes).
    void clear_ring_levels(int from, int to) {
        const uint32_t W = L.max + 1;
        const size_t stride = (size_t)slots * n;
        for (int x = std::max(from, 0); x <= to && x + 1 < (int)lv.size(); x++) {
            const uint32_t qb = lv[x], qe = lv[x + 1];
            if (qe == qb) continue;
            zero_entry_words_kernel<<<blocks_for_threads(std::min<size_t>(qe - qb, (size_t)1 << 20)),
                                      kBlock, 0, st>>>(ws.q.p, qb, qe,
                                                       ws.ring.p + (size_t)(x % W) * stride, n);
            launch_check();
        }
    }

    // The dependency sweep of a ring search: the breadth-first sweep, one
    // distance at a time from the last, with successors at distance D + len
    // (RingBits<false>). The ring starts empty and is rebuilt as the sweep
    // goes down (slot D mod W: distance D in, distance D + W out), so at
    // distance D it holds distances D + 1 .. D + W - 1; it is left empty.
    template <class C, bool kEdge, bool kQuirk = false>
    void backward_ring(C* coef) {
        const DeviceGraph& d = g.dev;
        const uint32_t W = L.max + 1;
        const size_t stride = (size_t)slots * n;
        const int D = (int)lv.size() - 2;  // last distance with entries
        clear_ring_levels((int)ring_top - (int)W + 1, D);
        const int split = opt.exact || kQuirk ? 0 : 1;
        const int low = kEdge ? 0 : 1;
        auto kernel = bfs_backward_kernel<C, kEdge, kQuirk, RingBits<false>>;
        for (int dl = D; dl >= low; dl--) {
            const uint32_t qb = lv[dl], qe = lv[dl + 1];
            const uint32_t hb = hv[dl], he = hv[dl + 1];
            const bool far = dl + (int)W <= D;
            const uint32_t ab = far ? lv[dl + W] : 0, ae = far ? lv[dl + W + 1] : 0;
            if (qb == qe && ab == ae) continue;
            const bool heavy = split && dl < D && he > hb;
            uint32_t gsz = group_size(qe - qb, resident_waves(kernel),
                                      (uint32_t)std::min(std::max(opt.sweep_group, 1), 64));
            if (!split && d.max_out * gsz > 0xFFFFFFFFull - 64) gsz = 1;
            const size_t tasks = (size_t)(qe - qb + gsz - 1) / gsz + (heavy ? hc[dl] : 0);
            const size_t waves = std::max<size_t>(tasks, (qe - qb + ae - ab + 63) / 64);
            if (heavy && ws.part.n < (size_t)hc[dl] * kLanes) {
                NETBC_HIP(hipStreamSynchronize(st));  // the old buffer may be in use
                ws.part.alloc((size_t)hc[dl] * kLanes * 2);
            }
            const uint32_t d0 = (uint32_t)dl % W;
            ws.sweep.mark(st);
            kernel<<<ticket_grid(resident_blocks(kernel, waves)), kBlock, 0, st>>>(
                ws.q.p, qb, qe, ab, ae, dl == D, 1, ws.hq.p, hb, he, ws.hpre.p, hc[dl], split,
                d.heavy_flag.p, ws.part.p, d.out_off.p, d.out_idx.p,
                RingBits<false>{ws.ring.p, stride, out_lengths<IntPath>(), d0, W},
                ws.ring.p + (size_t)d0 * stride, ws.succ.p, ws.sig.p, coef, ws.ctrl.p, gsz, n,
                ws.eacc.p, d.arcs, target, npr, dl > 0 ? bsum : nullptr, nullptr);
            launch_check();
            if (heavy) {
                bfs_combine_kernel<C><<<resident_blocks(bfs_combine_kernel<C>, he - hb), kBlock, 0,
                                        st>>>(ws.q.p, ws.hq.p, hb, he, ws.hpre.p, hc[dl],
                                              ws.part.p, d.out_off.p, ws.sig.p, coef, ws.ctrl.p,
                                              n, target, dl > 0 ? bsum : nullptr);
                launch_check();
            }
            ws.sweep.mark(st);
        }
        clear_ring_levels(low, std::min(D, low + (int)W - 1));
    }

    void run_dijkstra(int64_t item0, DjkState& dst) {
        const DeviceGraph& d = g.dev;
        reset_slots();
        ws.sweep.mark(st);
        dijkstra_kernel<<<slots, kLanes, 0, st>>>(
            d_list, item0, k, slots, n, d.out_off.p, d.out_idx.p, L.out_real.p, dst, ws.vis.p,
            ws.sig.p, ws.ctrl.p, edge ? ws.eacc.p : nullptr, d.arcs, target);
        launch_check();
        ws.sweep.mark(st);
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
    ws.hq.release();
    ws.hq.alloc(nq);
    ws.hpre.release();
    ws.hpre.alloc(nq + 1);
    ws.perm.release();
    ws.perm.alloc(nq);
    ws.qcap = nq;
}

void ensure_workspace(Workspace& ws, const Graph& g, const Lengths& L, int slots, int precision,
                      size_t qcap, bool ring = false) {
    size_t n = std::max<uint64_t>(g.n, 1);
    size_t nm = (size_t)slots * n;
    ws.slots = slots;
    ws.n = g.n;
    ws.vis.alloc(nm);
    if (ws.sig.alloc(nm * kLanes)) ws.rows_dirty = true;  // rows must start at zero
    if (L.kind == kUniform) {
        ws.cand.alloc(nm);
        ws.succ.alloc(nm);
        ws.flanes.alloc((size_t)2 * slots);
        if (ws.levels.alloc(nm * kLevelBufs)) ws.zero_words = ws.levels.n;
        ws.touched.alloc(nm);
        ws.bsum.alloc(nm);
    }
    const bool per_thread = L.kind == kReal && !real_bucketed(L, g.n);
    if (!per_thread) {
        if (precision) ws.coef32.alloc(nm * kLanes);
        else ws.coef64.alloc(nm * kLanes);
        if (qcap > ws.qcap || !ws.q.p) {
            ws.q.release();
            ws.q.alloc(qcap);
            ws.hq.release();
            ws.hq.alloc(qcap);
            ws.hpre.release();
            ws.hpre.alloc(qcap + 1);
            ws.perm.release();
            ws.perm.alloc(qcap);
            ws.qcap = qcap;
        }
    }
    if (ring) {
        const size_t W = (size_t)L.max + 1;
        ws.succ.alloc(nm);
        ws.bsum.alloc(nm);
        const bool fresh_ring = ws.ring.alloc(nm * W);
        if (ws.cring.alloc(nm * W) || fresh_ring) ws.ring_dirty = true;
        if (ws.lvl_cap == 0) {
            ws.lvl_cap = 4096;
            ws.lvl_q.alloc(ws.lvl_cap);
            ws.lvl_h.alloc(ws.lvl_cap);
            ws.hc_lvl.alloc(ws.lvl_cap);
        }
    } else if (L.kind == kInteger || (L.kind == kReal && !per_thread)) {
        if (L.kind == kInteger) ws.dist.alloc(nm * kLanes);
        else ws.rdist.alloc(nm * kLanes);
        ws.inb.alloc(nm * buckets_for(L).words);
        ws.dlist.alloc(nm);
    }
    ws.acc.alloc(n);
    ws.gcur.alloc(n);
    ws.ctrl.alloc(1);
    ws.slot_cnt.alloc((size_t)std::max(slots, 1));
}

// Weighted distances from every node to the first kWeightedLandmarks
// landmarks by Bellman-Ford rounds (until no distance falls), written as
// bytes (distance / unit, capped at 255) into the landmark table `dist`.
template <class T>
void landmark_distances(const DeviceGraph& dg, const T* len, int lanes, double unit,
                        uint8_t* dist, hipStream_t st) {
    const size_t n = dg.n;
    const int wl = std::min(lanes, kWeightedLandmarks);
    DevBuf<T> wd;
    DevBuf<uint32_t> flag;
    wd.alloc(n * kWeightedLandmarks);
    flag.alloc(1);
    std::vector<T> init(n * kWeightedLandmarks,
                        std::is_same_v<T, uint32_t> ? (T)kNone : (T)INFINITY);
    for (int j = 0; j < wl; j++) init[(size_t)dg.landmarks[j] * kWeightedLandmarks + j] = (T)0;
    NETBC_HIP(hipMemcpyAsync(wd.p, init.data(), init.size() * sizeof(T), hipMemcpyHostToDevice,
                             st));
    const unsigned rb = resident_blocks(landmark_relax_kernel<T>, n);
    for (size_t round = 0; round < n; round++) {
        NETBC_HIP(hipMemsetAsync(flag.p, 0, 4, st));
        landmark_relax_kernel<T><<<rb, kBlock, 0, st>>>(dg.out_off.p, dg.out_idx.p, len, wd.p, n,
                                                         flag.p);
        NETBC_HIP(hipGetLastError());
        uint32_t changed = 0;
        NETBC_HIP(hipMemcpyAsync(&changed, flag.p, 4, hipMemcpyDeviceToHost, st));
        NETBC_HIP(hipStreamSynchronize(st));  // also keeps `init` alive for its copy
        if (!changed) break;
    }
    landmark_bytes_kernel<T><<<blocks_for_threads(std::min<size_t>(n, (size_t)1 << 22)), kBlock, 0,
                               st>>>(wd.p, n, unit, dist);
    NETBC_HIP(hipGetLastError());
    NETBC_HIP(hipStreamSynchronize(st));
}

// Orders the k sources of a free-order run (list d_list, nullptr: 0..k-1)
// so that sources whose searches look alike share batches: three stable
// sorts, by second out-neighbour, by first out-neighbour, and by the
// distances to the graph's landmarks (capped at 15), the last one deciding
// first. Sources in one batch then reach most nodes on the same level, so
// the 64-lane rows carry more lanes per entry. With integer lengths (and
// bucketed real ones, in units of the shortest arc) the last sort uses the
// weighted distances to the first 8 landmarks instead (capped at 255):
// sources at equal lengths from the hubs
// settle most nodes in the same buckets (on Barabasi-Albert n = 200,000
// with lengths 1-16, 9-29 lanes per settled entry instead of 4-5 by hop
// counts). Writes the list to ws.list2.
void order_sources(Graph& g, Workspace& ws, const int64_t* d_list, int64_t k, hipStream_t st,
                   const Lengths* L = nullptr) {
    DeviceGraph& dg = g.dev;
    const size_t n = g.n;
    const int lanes = (int)dg.landmarks.size();
    if (!dg.lm_hops_ready) {
        DevBuf<uint8_t>& dist = dg.lm_hops;
        dist.alloc(n * kLandmarks);
        NETBC_HIP(hipMemsetAsync(dist.p, 15, n * kLandmarks, st));
        DevBuf<uint32_t> cur, vis, nxt;
        cur.alloc(n);
        vis.alloc(n);
        nxt.alloc(n);
        NETBC_HIP(hipMemsetAsync(cur.p, 0, n * 4, st));
        NETBC_HIP(hipMemsetAsync(nxt.p, 0, n * 4, st));
        if (lanes > 0) {
            // level 0: the landmarks themselves
            std::vector<uint32_t> seed(n, 0);
            for (int j = 0; j < lanes; j++) seed[dg.landmarks[j]] |= 1u << j;
            NETBC_HIP(hipMemcpyAsync(cur.p, seed.data(), n * 4, hipMemcpyHostToDevice, st));
            NETBC_HIP(hipMemcpyAsync(vis.p, seed.data(), n * 4, hipMemcpyHostToDevice, st));
            for (int j = 0; j < lanes; j++)
                NETBC_HIP(hipMemsetAsync(dist.p + (size_t)dg.landmarks[j] * kLandmarks + j, 0, 1,
                                         st));
            const unsigned offer_blocks = resident_blocks(landmark_offer_kernel, (n + 63) / 64);
            const unsigned settle_blocks =
                blocks_for_threads(std::min<size_t>(n, (size_t)1 << 22));
            for (int level = 1; level < 15; level++) {
                landmark_offer_kernel<<<offer_blocks, kBlock, 0, st>>>(
                    dg.in_off_p(), dg.in_idx_p(), cur.p, vis.p, nxt.p, n);
                NETBC_HIP(hipGetLastError());
                landmark_settle_kernel<<<settle_blocks, kBlock, 0, st>>>(
                    cur.p, vis.p, nxt.p, dist.p, level, lanes, n);
                NETBC_HIP(hipGetLastError());
            }
            // the seed buffer must outlive the copies
            NETBC_HIP(hipStreamSynchronize(st));
        }
        dg.lm_hops_ready = true;
    }
    const uint8_t* table = dg.lm_hops.p;
    // weighted runs: this run's weighted distances in the first columns of
    // a copy of the table
    DevBuf<uint8_t> wdist;
    const bool weighted =
        L && lanes > 0 && (L->kind == kInteger || (L->kind == kReal && real_bucketed(*L, n)));
    if (weighted) {
        wdist.alloc(n * kLandmarks);
        NETBC_HIP(hipMemcpyAsync(wdist.p, dg.lm_hops.p, n * kLandmarks, hipMemcpyDeviceToDevice,
                                 st));
        if (L->kind == kInteger)
            landmark_distances<uint32_t>(dg, L->out_len.p, lanes, 1.0, wdist.p, st);
        else
            landmark_distances<double>(dg, L->out_real.p, lanes, L->rmin, wdist.p, st);
        table = wdist.p;
    }
    DevBuf<uint64_t> k0, k1;
    DevBuf<uint32_t> v0, v1;
    k0.alloc(k);
    k1.alloc(k);
    v0.alloc(k);
    v1.alloc(k);
    iota_kernel<<<blocks_for_threads(k), kBlock, 0, st>>>(v0.p, k);
    NETBC_HIP(hipGetLastError());
    uint64_t* key = k0.p;
    uint64_t* key_tmp = k1.p;
    uint32_t* perm = v0.p;
    uint32_t* perm_tmp = v1.p;
    const int b = node_bits(n + 1);
    for (int which = 0; which < 3; which++) {
        const int kind = which == 2 && weighted ? 3 : which;
        source_key_kernel<<<blocks_for_threads(k), kBlock, 0, st>>>(
            d_list, k, perm, dg.out_off.p, dg.out_idx.p, table, (uint32_t)n, kind, key);
        NETBC_HIP(hipGetLastError());
        const int bits = which < 2 ? b : weighted ? 8 * kWeightedLandmarks : 4 * kLandmarks;
        if (sort_pairs_u64(key, perm, key_tmp, perm_tmp, k, bits, st)) {
            std::swap(key, key_tmp);
            std::swap(perm, perm_tmp);
        }
    }
    ws.list2.alloc(k);
    gather_sources_kernel<<<blocks_for_threads(k), kBlock, 0, st>>>(d_list, k, perm, ws.list2.p);
    NETBC_HIP(hipGetLastError());
    NETBC_HIP(hipStreamSynchronize(st));
}

}  // namespace

void run_batched(Graph& g, const Lengths& L, const int64_t* sources, int64_t k,
                 const RunOptions& opt_in, double* d_bc, RunDiag& diag, double* d_eb,
                 const SubsetRun* sub) {
    const RunOptions opt = tuned_options(g, L, opt_in);
    hipStream_t st = g.stream;
    size_t n = g.n;
    const bool edge = d_eb != nullptr;
    const size_t arcs = g.dev.arcs;
    diag.path_kind = L.kind;
    if (n == 0) return;
    NETBC_HIP(hipMemsetAsync(d_bc, 0, n * sizeof(double), st));
    if (edge && g.m) NETBC_HIP(hipMemsetAsync(d_eb, 0, g.m * sizeof(double), st));
    if (k == 0) return;
    Workspace& ws = get_workspace(g);
    // hub tables cut for this run's thresholds; rows of new hubs start at zero
    if (ensure_hub_tables(g, (uint32_t)opt.hub_in, (uint32_t)opt.hub_chunk)) ws.rows_dirty = true;
    const bool ring = ring_path(g, L);
    int slots = choose_slots(g, L, k, opt, edge, ring);
    diag.slots = slots;
    size_t qcap = std::max<size_t>(queue_capacity(slots, n), ws.qcap);
    ensure_workspace(ws, g, L, slots, opt.precision, qcap, ring);
    NETBC_HIP(hipMemsetAsync(ws.acc.p, 0, n * sizeof(double), st));
    NETBC_HIP(hipMemsetAsync(ws.gcur.p, 0, n * sizeof(int64_t), st));
    const int64_t* d_list = nullptr;
    if (sources) {
        ws.list.alloc(k);
        NETBC_HIP(hipMemcpyAsync(ws.list.p, sources, k * sizeof(int64_t), hipMemcpyHostToDevice,
                                 st));
        d_list = ws.list.p;
    }
    // real lengths with very unequal arcs: the reference's Dijkstra per thread
    const bool per_thread = L.kind == kReal && !real_bucketed(L, n);
    if (per_thread && sub && sub->tie_exact)
        throw Error("exact ties on real lengths whose longest arc exceeds 2,047 times the "
                    "shortest are not available on the device path");
    const bool free_order = !opt.exact && k >= kReorderMin && !per_thread;
    if (free_order) {
        // Batch sources whose searches look alike (order_sources). Items are
        // summed in the new order, which changes only the association of
        // the cross-source sum.
        order_sources(g, ws, d_list, k, st, &L);
        d_list = ws.list2.p;
    }
    NETBC_HIP(hipMemsetAsync(ws.ctrl.p, 0, sizeof(Ctrl), st));
    ws.clean_rows(st);
    ws.rows_dirty = true;
    NETBC_HIP(hipStreamSynchronize(st));
    ws.sweep.collect();
    ws.sweep.ms = 0.0;
    ws.sweep.launches = 0;
    DevBuf<double> djk_buf;
    DevBuf<uint32_t> djk_u32;
    DjkState djk{};
    if (per_thread) {
        size_t T = (size_t)slots * kLanes * n;
        djk_buf.alloc(4 * T);
        djk_u32.alloc(4 * T);
        djk = DjkState{djk_buf.p, djk_buf.p + T, djk_buf.p + 2 * T, djk_u32.p, djk_u32.p + T,
                       djk_buf.p + 3 * T, djk_u32.p + 2 * T, djk_u32.p + 3 * T};
    }
    int rows = 0;  // edge credit rows (Workspace::eacc)
    if (edge) {
        rows = per_thread ? slots * kLanes : slots;
        ws.eacc.alloc((size_t)rows * arcs);
        NETBC_HIP(hipMemsetAsync(ws.eacc.p, 0, (size_t)rows * arcs * sizeof(double), st));
    }
    Engine eng{g, L, ws, opt, diag, st, n, slots, d_list, k, edge};
    eng.ring = ring;
    // dependency sums from the sweep
    const bool fold = free_order && (L.kind == kUniform || ring);
    if (fold) {
        NETBC_HIP(hipMemsetAsync(ws.bsum.p, 0, (size_t)slots * n * sizeof(double), st));
        eng.bsum = ws.bsum.p;
    }
    const uint8_t* d_target = sub ? sub->d_target : nullptr;
    const bool endpoints = sub && sub->endpoints;
    eng.target = d_target;
    eng.tie_exact = sub && sub->tie_exact;
    if (sub && sub->edge_split_nx) {
        if (!edge) throw Error("NetworkX's edge split needs an edge run");
        if (per_thread)
            throw Error("NetworkX's subset edge split on real lengths whose longest arc exceeds "
                        "2,047 times the shortest is not available on the device path");
        eng.quirk = true;
        ws.npr.alloc((size_t)slots * n * kLanes);
        NETBC_HIP(hipMemsetAsync(ws.npr.p, 0, (size_t)slots * n * kLanes * 4, st));
        eng.npr = ws.npr.p;
    }
    DevBuf<unsigned long long> ep, lane_count;  // endpoint counts
    if (endpoints) {
        ep.alloc(n);
        lane_count.alloc((size_t)slots * kLanes);
        NETBC_HIP(hipMemsetAsync(ep.p, 0, n * sizeof(unsigned long long), st));
        NETBC_HIP(hipMemsetAsync(lane_count.p, 0, (size_t)slots * kLanes * 8, st));
    }
    int64_t gsize = (k + 63) / 64;
    NETBC_HIP(hipEventRecord(ws.ev0, st));
    for (int64_t item0 = 0; item0 < k; item0 += (int64_t)slots * kLanes) {
        if (per_thread) {
            eng.run_dijkstra(item0, djk);
        } else {
            for (;;) {
                bool ok = L.kind == kUniform   ? eng.run_bfs(item0)
                          : ring               ? eng.run_ring(item0)
                          : L.kind == kInteger ? eng.run_bucketed<IntPath>(item0)
                          : eng.tie_exact      ? eng.run_bucketed<RealPathExact>(item0)
                                               : eng.run_bucketed<RealPath>(item0);
                if (ok) break;
                // Level queue overflow: grow and repeat the super-batch from
                // clean rows.
                eng.clear_rows();
                grow_queue(ws);
            }
            if (opt.precision) eng.backward(ws.coef32.p);
            else eng.backward(ws.coef64.p);
        }
        if (endpoints) {
            // before the row sums, which leave vis alone but come last
            endpoint_nodes_kernel<<<blocks_for_threads(std::min<size_t>(n, (size_t)1 << 22)),
                                    kBlock, 0, st>>>(ws.vis.p, slots, n, d_target, ep.p);
            endpoint_lanes_kernel<<<resident_blocks(endpoint_lanes_kernel,
                                                    (n + 63) / 64 * (size_t)slots),
                                    kBlock, 0, st>>>(ws.vis.p, slots, n, d_target, lane_count.p);
            endpoint_sources_kernel<<<blocks_for_threads((size_t)slots * kLanes), kBlock, 0,
                                      st>>>(d_list, item0, k, slots, d_target, lane_count.p,
                                            ep.p);
            NETBC_HIP(hipGetLastError());
        }
        if (fold) {
            // slots holding sources of this super-batch
            const int active = (int)std::min<int64_t>(slots, (k - item0 + kLanes - 1) / kLanes);
            fold_slot_sums_kernel<<<blocks_for_threads(n), kBlock, 0, st>>>(
                ws.vis.p, ws.bsum.p, active, n, d_bc, g.dev.out_off.p, ws.ctrl.p);
            if (g.dev.hubs && L.kind == kUniform) {  // hub rows hold atomic partial sums
                NETBC_HIP(hipGetLastError());
                const size_t rows = (size_t)slots * g.dev.hubs;
                zero_hub_rows_kernel<<<blocks_for_threads(std::min<size_t>(rows * 64,
                                                                           (size_t)1 << 22)),
                                       kBlock, 0, st>>>(ws.vis.p, ws.sig.p, g.dev.hub_list.p,
                                                        g.dev.hubs, slots, n);
            }
        } else if (free_order) {
            const uint32_t nsrc = eng.lv.size() > 1 ? eng.lv[1] : 0;
            if (nsrc) {
                zero_sources_kernel<<<blocks_for_threads(nsrc), kBlock, 0, st>>>(ws.q.p, nsrc,
                                                                                ws.sig.p, n);
                NETBC_HIP(hipGetLastError());
            }
            accumulate_rows_kernel<<<resident_blocks(accumulate_rows_kernel, n), kBlock, 0,
                                     st>>>(ws.vis.p, ws.sig.p, slots, n, item0, k, d_bc,
                                           g.dev.out_off.p, ws.ctrl.p);
            if (g.dev.hubs) {
                NETBC_HIP(hipGetLastError());
                const size_t rows = (size_t)slots * g.dev.hubs;
                zero_hub_rows_kernel<<<blocks_for_threads(std::min<size_t>(rows * 64,
                                                                           (size_t)1 << 22)),
                                       kBlock, 0, st>>>(ws.vis.p, ws.sig.p, g.dev.hub_list.p,
                                                        g.dev.hubs, slots, n);
            }
        } else if (std::min<int64_t>(k - item0, (int64_t)slots * kLanes) >= 16 * gsize) {
            // many groups per super-batch: a wavefront per node, a lane per group
            accumulate_groups_kernel<<<resident_blocks(accumulate_groups_kernel, n), kBlock, 0,
                                       st>>>(ws.vis.p, ws.sig.p, slots, n, item0, k, gsize,
                                             d_list, ws.acc.p, d_bc, ws.gcur.p, g.dev.out_off.p,
                                             ws.ctrl.p);
        } else
            accumulate_kernel<<<blocks_for_threads(n), kBlock, 0, st>>>(
                ws.vis.p, ws.sig.p, slots, n, item0, k, gsize, d_list, ws.acc.p, d_bc, ws.gcur.p,
                g.dev.out_off.p, ws.ctrl.p);
        NETBC_HIP(hipGetLastError());
    }
    finish_kernel<<<blocks_for_threads(n), kBlock, 0, st>>>(d_bc, ws.acc.p, n);
    NETBC_HIP(hipGetLastError());
    if (endpoints) {
        add_counts_kernel<<<blocks_for_threads(std::min<size_t>(n, (size_t)1 << 22)), kBlock, 0,
                            st>>>(d_bc, ep.p, n);
        NETBC_HIP(hipGetLastError());
    }
    if (edge && arcs) {
        fold_arcs_kernel<<<blocks_for_threads(std::min<size_t>(arcs, (size_t)1 << 22)), kBlock, 0,
                           st>>>(ws.eacc.p, rows, arcs, g.dev.arc_edge.p, d_eb);
        NETBC_HIP(hipGetLastError());
    }
    NETBC_HIP(hipEventRecord(ws.ev1, st));
    read_ctrl(ws, st);
    ws.rows_dirty = false;
    ws.sweep.collect();
    float ms = 0.0f;
    NETBC_HIP(hipEventElapsedTime(&ms, ws.ev0, ws.ev1));
    diag.kernel_ms += ms;
    diag.dominant_ms += ws.sweep.ms;
    diag.dominant_launches += ws.sweep.launches;
    Ctrl& c = *ws.host_ctrl;
    double mx;
    unsigned long long bits = c.max_sigma_bits;
    memcpy(&mx, &bits, 8);
    diag.max_sigma = std::max(diag.max_sigma, mx);
    diag.sigma_overflow |= c.sig_overflow ? 1 : 0;
    diag.reached_nodes += (int64_t)spread_sum(c.reached_nodes);
    diag.reached_arcs += (int64_t)spread_sum(c.reached_arcs);
    for (int i = 0; i < 16; i++) diag.work[i] += (int64_t)c.work[i];
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

void scale_flagged(Graph& g, double* d_vals, size_t count, const uint8_t* d_flag,
                   double f_flagged, double f_other) {
    if (count == 0) return;
    scale_flagged_kernel<<<blocks_for_threads(std::min<size_t>(count, (size_t)1 << 22)), kBlock,
                           0, g.stream>>>(d_vals, count, d_flag, f_flagged, f_other);
    NETBC_HIP(hipGetLastError());
}

void finish_edge_scores(Graph& g, bool normalized, double* d_eb, double* out) {
    hipStream_t st = g.stream;
    const uint64_t n = g.n, m = g.m;
    if (m == 0) return;
    // reach.rs:207-224
    double scale = g.directed ? 1.0 : 0.5;
    if (normalized) {
        const double pairs = (double)n * ((double)n - 1.0);
        scale = n < 2 ? NAN : g.directed ? scale / pairs : scale * 2.0 / pairs;
    }
    if (scale != 1.0) {
        scale_kernel<<<blocks_for_threads(m), kBlock, 0, st>>>(d_eb, m, scale);
        NETBC_HIP(hipGetLastError());
    }
    NETBC_HIP(hipMemcpyAsync(out, d_eb, m * sizeof(double), hipMemcpyDeviceToHost, st));
    NETBC_HIP(hipStreamSynchronize(st));
}

// Direction factor of reachability-only breadth-first searches when the
// caller sets none: their bottom-up steps stop at the first parents and
// take no path counts, so they pay off earlier than in counting searches
// (R-MAT 18 closeness: 0.285 s at 1, 0.209 s at 2, 0.190 s at 4, 0.188 s at
// 8, 0.195 s at 16).
constexpr double kReachAlpha = 4.0;

void run_closeness(Graph& g, const Lengths& L, const RunOptions& opt_in, const ClosenessRun& cr,
                   double* out, RunDiag& diag) {
    RunOptions opt0 = opt_in;
    if (!(opt0.alpha > 0.0)) opt0.alpha = kReachAlpha;
    const RunOptions opt = tuned_options(g, L, opt0);
    hipStream_t st = g.stream;
    const size_t n = g.n;
    diag.path_kind = L.kind;
    if (n == 0) return;
    const int64_t k = (int64_t)n;
    Workspace& ws = get_workspace(g);
    // hub tables cut for this run's thresholds; rows of new hubs start at zero
    if (ensure_hub_tables(g, (uint32_t)opt.hub_in, (uint32_t)opt.hub_chunk)) ws.rows_dirty = true;
    const bool per_thread = L.kind == kReal && !real_bucketed(L, n);
    const int slots = choose_slots(g, L, k, opt, false);
    diag.slots = slots;
    ensure_workspace(ws, g, L, slots, opt.precision,
                     std::max<size_t>(queue_capacity(slots, n), ws.qcap));
    const size_t S = (size_t)slots * kLanes;
    DevBuf<uint64_t> pbuf, lbuf;  // profiles per node; per lane + level counts
    pbuf.alloc(5 * n);
    lbuf.alloc(6 * S);
    NETBC_HIP(hipMemsetAsync(pbuf.p, 0, 5 * n * 8, st));
    NETBC_HIP(hipMemsetAsync(lbuf.p, 0, 6 * S * 8, st));
    const DevProfile prof = profile_view(pbuf.p, n);
    const DevProfile lanes = profile_view(lbuf.p, S);
    const int64_t* d_list = nullptr;
    // Every node is a source once, so the batching order changes nothing
    // in the results (each profile is summed within its own search).
    if (!opt.exact && k >= kReorderMin && !per_thread) {
        order_sources(g, ws, nullptr, k, st, &L);
        d_list = ws.list2.p;
    }
    NETBC_HIP(hipMemsetAsync(ws.ctrl.p, 0, sizeof(Ctrl), st));
    ws.clean_rows(st);
    ws.rows_dirty = true;
    NETBC_HIP(hipStreamSynchronize(st));
    ws.sweep.collect();
    ws.sweep.ms = 0.0;
    ws.sweep.launches = 0;
    DevBuf<double> djk_buf;
    DevBuf<uint32_t> djk_u32;
    DjkState djk{};
    if (per_thread) {
        const size_t T = S * n;
        djk_buf.alloc(2 * T);
        djk_u32.alloc(3 * T);
        djk = DjkState{djk_buf.p, nullptr, nullptr, djk_u32.p, nullptr, djk_buf.p + T,
                       djk_u32.p + T, djk_u32.p + 2 * T};
    }
    const size_t chunks = (n + kProfileChunk - 1) / kProfileChunk;
    DevBuf<uint32_t> pc;
    DevBuf<unsigned long long> pit;
    DevBuf<double> pft, ph, pl;
    if (!per_thread && L.kind != kUniform) {
        pc.alloc(chunks * S);
        pit.alloc(chunks * S);
        pft.alloc(chunks * S);
        ph.alloc(chunks * S);
        pl.alloc(chunks * S);
    }
    NETBC_HIP(hipStreamSynchronize(st));
    Engine eng{g, L, ws, opt, diag, st, n, slots, d_list, k, false};
    eng.reach_only = true;
    if (L.kind == kUniform) {
        eng.lane_buf = lbuf.p;
        eng.lane_count = (unsigned long long*)lbuf.p + 5 * S;
    }
    const DeviceGraph& dg = g.dev;
    NETBC_HIP(hipEventRecord(ws.ev0, st));
    for (int64_t item0 = 0; item0 < k; item0 += (int64_t)S) {
        if (per_thread) {
            dijkstra_profile_kernel<<<slots, kLanes, 0, st>>>(d_list, item0, k, slots, n,
                                                             dg.out_off.p, dg.out_idx.p,
                                                             L.out_real.p, djk, lanes);
            NETBC_HIP(hipGetLastError());
        } else {
            for (;;) {
                const bool ok = L.kind == kUniform   ? eng.run_bfs(item0)
                                : L.kind == kInteger ? eng.run_bucketed<IntPath>(item0)
                                                     : eng.run_bucketed<RealPath>(item0);
                if (ok) break;
                eng.clear_rows();
                grow_queue(ws);
            }
            if (L.kind == kInteger) eng.dist_profile<IntPath>(lanes, pc.p, pit.p, pft.p, ph.p, pl.p);
            else if (L.kind == kReal)
                eng.dist_profile<RealPath>(lanes, pc.p, pit.p, pft.p, ph.p, pl.p);
            // only the sources' own path counts were written
            const uint32_t nsrc = eng.lv[1];
            if (nsrc) {
                zero_sources_kernel<<<blocks_for_threads(nsrc), kBlock, 0, st>>>(ws.q.p, nsrc,
                                                                                ws.sig.p, n);
                NETBC_HIP(hipGetLastError());
            }
        }
        commit_profile_kernel<<<blocks_for_threads(S), kBlock, 0, st>>>(lanes, slots, d_list,
                                                                        item0, k, prof);
        NETBC_HIP(hipGetLastError());
    }
    // values, closeness.rs:131-209
    DevBuf<unsigned long long> longest;
    longest.alloc(1);
    NETBC_HIP(hipMemsetAsync(longest.p, 0, 8, st));
    if (!cr.harmonic && cr.unreachable == 3) {
        max_bits_kernel<<<blocks_for_threads(std::min<size_t>(n, (size_t)1 << 20)), kBlock, 0,
                          st>>>(prof.longest, n, longest.p);
        NETBC_HIP(hipGetLastError());
    }
    DevBuf<double> vals;
    vals.alloc(n);
    const unsigned vb = blocks_for_threads(std::min<size_t>(n, (size_t)1 << 22));
    if (cr.d_types) {
        if (L.kind != kUniform) throw Error("two-mode closeness counts steps");
        twomode_closeness_kernel<<<vb, kBlock, 0, st>>>(prof, n, cr.d_types, (double)cr.first,
                                                        (double)cr.second, cr.normalized ? 1 : 0,
                                                        cr.unreachable == 1 ? 1 : 0, vals.p);
    } else {
        closeness_kernel<<<vb, kBlock, 0, st>>>(prof, n, L.kind == kReal ? 1 : 0, cr.unit,
                                                cr.harmonic ? 1 : 0, cr.normalized ? 1 : 0,
                                                cr.unreachable, longest.p, vals.p);
    }
    NETBC_HIP(hipGetLastError());
    NETBC_HIP(hipEventRecord(ws.ev1, st));
    std::vector<unsigned long long> reached(n);
    NETBC_HIP(hipMemcpyAsync(out, vals.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
    NETBC_HIP(hipMemcpyAsync(reached.data(), prof.reached, n * 8, hipMemcpyDeviceToHost, st));
    NETBC_HIP(hipStreamSynchronize(st));
    ws.rows_dirty = false;
    float ms = 0.0f;
    NETBC_HIP(hipEventElapsedTime(&ms, ws.ev0, ws.ev1));
    diag.kernel_ms += ms;
    for (size_t v = 0; v < n; v++) diag.reached_nodes += (int64_t)reached[v];
    read_ctrl(ws, st);
    ws.sweep.collect();
    diag.dominant_ms += ws.sweep.ms;
    diag.dominant_launches += ws.sweep.launches;
    diag.dominant_bytes += (int64_t)(spread_sum(ws.host_ctrl->reach_bytes) + eng.stale_bytes);
    double mx;
    const unsigned long long bits = ws.host_ctrl->max_sigma_bits;
    memcpy(&mx, &bits, 8);
    diag.max_sigma = std::max(diag.max_sigma, mx);
}

void degree_scores(Graph& g, int mode, bool loops, double* d_out) {
    hipStream_t st = g.stream;
    NETBC_HIP(hipMemsetAsync(d_out, 0, std::max<uint64_t>(g.n, 1) * sizeof(double), st));
    if (g.m == 0) return;
    DevBuf<uint32_t> ends;
    ends.alloc(2 * g.m);
    NETBC_HIP(hipMemcpyAsync(ends.p, g.src.data(), g.m * 4, hipMemcpyHostToDevice, st));
    NETBC_HIP(hipMemcpyAsync(ends.p + g.m, g.tgt.data(), g.m * 4, hipMemcpyHostToDevice, st));
    degree_kernel<<<blocks_for_threads(std::min<size_t>(g.m, (size_t)1 << 22)), kBlock, 0, st>>>(
        ends.p, ends.p + g.m, g.m, g.directed ? 1 : 0, mode, loops ? 1 : 0, d_out);
    NETBC_HIP(hipGetLastError());
    NETBC_HIP(hipStreamSynchronize(st));  // the edge copies must finish first
}

void divide_by_type(Graph& g, double* d_vals, const uint8_t* d_types, double div0, double div1) {
    if (g.n == 0) return;
    divide_by_type_kernel<<<blocks_for_threads(std::min<size_t>(g.n, (size_t)1 << 22)), kBlock, 0,
                            g.stream>>>(d_vals, g.n, d_types, div0, div1);
    NETBC_HIP(hipGetLastError());
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
    RunOptions o = tuned_options(g, L, opt);
    // hub tables cut for this run's thresholds; rows of new hubs start at zero
    if (ensure_hub_tables(g, (uint32_t)o.hub_in, (uint32_t)o.hub_chunk)) ws.rows_dirty = true;
    const bool ring = ring_path(g, L);
    int slots = choose_slots(g, L, k, o, false, ring);
    size_t qcap = std::max<size_t>(queue_capacity(slots, n), ws.qcap);
    ensure_workspace(ws, g, L, slots, o.precision, qcap, ring);
    ws.list.alloc(k);
    NETBC_HIP(hipMemcpyAsync(ws.list.p, sources, k * sizeof(int64_t), hipMemcpyHostToDevice, st));
    NETBC_HIP(hipMemsetAsync(ws.ctrl.p, 0, sizeof(Ctrl), st));
    ws.clean_rows(st);
    ws.rows_dirty = true;
    RunDiag diag;
    Engine eng{g, L, ws, o, diag, st, n, slots, ws.list.p, k};
    eng.ring = ring;
    std::vector<double> rows((size_t)slots * n * kLanes);
    std::vector<uint32_t> drows;
    std::vector<uint64_t> vis((size_t)slots * n);
    for (int64_t item0 = 0; item0 < k; item0 += (int64_t)slots * kLanes) {
        for (;;) {
            bool ok = L.kind == kUniform ? eng.run_bfs(item0)
                      : ring               ? eng.run_ring(item0)
                                           : eng.run_bucketed<IntPath>(item0);
            if (ok) break;
            eng.clear_rows();
            grow_queue(ws);
        }
        int64_t cnt = std::min<int64_t>((int64_t)slots * kLanes, k - item0);
        NETBC_HIP(hipMemcpyAsync(rows.data(), ws.sig.p, rows.size() * 8, hipMemcpyDeviceToHost, st));
        NETBC_HIP(hipMemcpyAsync(vis.data(), ws.vis.p, vis.size() * 8, hipMemcpyDeviceToHost, st));
        if (L.kind == kInteger && !ring) {
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
                if (r && L.kind == kInteger && !ring)
                    dist[o2 + v] = drows[((size_t)b * n + v) * kLanes + l];
                delta[o2 + v] = 0.0;
            }
        }
        if (L.kind == kUniform || ring) {  // the level of an entry is its distance
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
        if (o.precision) eng.backward(ws.coef32.p);
        else eng.backward(ws.coef64.p);
        NETBC_HIP(hipMemcpyAsync(rows.data(), ws.sig.p, rows.size() * 8, hipMemcpyDeviceToHost, st));
        eng.clear_rows();
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
    ws.rows_dirty = false;
}

}  // namespace netbc
