/*
 * Seeded workload generators for betweenness runs (host C).
 *
 * The random stream is Philox4x32-10 named by (seed, stream index), as in
 * knappnet's crates/engine/src/rng.rs:22-110, so an edge list depends only
 * on its seed. G(n, m), Watts-Strogatz and Barabasi-Albert follow
 * crates/engine/src/generate.rs (gnm_one 255-290, small_world_one 533-581,
 * preferential_one 602-631); the reference has no R-MAT generator, so
 * gen_rmat is this library's own definition (documented at the function).
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- Philox4x32-10: rng.rs:10-34 --------------------------------------- */

static inline void philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; round++) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void gen_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox(ctr, key, out);
}

/* Stream, rng.rs:37-110. */
typedef struct {
    uint32_t key[2];
    uint64_t stream, block;
    uint32_t buf[4];
    int used;
} Stream;

static void st_init(Stream *s, uint64_t seed, uint64_t index) {
    s->key[0] = (uint32_t)seed;
    s->key[1] = (uint32_t)(seed >> 32);
    s->stream = index;
    s->block = 0;
    s->used = 4;
}

static inline uint32_t st_u32(Stream *s) {
    if (s->used == 4) {
        uint32_t ctr[4] = {(uint32_t)s->block, (uint32_t)(s->block >> 32), (uint32_t)s->stream,
                           (uint32_t)(s->stream >> 32)};
        philox(ctr, s->key, s->buf);
        s->block++;
        s->used = 0;
    }
    return s->buf[s->used++];
}

static inline uint64_t st_u64(Stream *s) {
    uint64_t hi = st_u32(s);
    return hi << 32 | st_u32(s);
}

static inline uint64_t st_below(Stream *s, uint64_t n) {
    uint64_t zone = UINT64_MAX - (UINT64_MAX % n + 1) % n;
    for (;;) {
        uint64_t x = st_u64(s);
        if (x <= zone) return x % n;
    }
}

static inline double st_uniform(Stream *s) {
    return (double)(st_u64(s) >> 11) * (1.0 / 9007199254740992.0);
}

/* The k-th uniform of Stream(seed, index), computed directly from the
 * counter (uniform k uses u32 words 2k and 2k+1 of the stream). */
static inline double uniform_at(const uint32_t key[2], uint64_t index, uint64_t k) {
    uint64_t blk = k >> 1;
    uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)index,
                       (uint32_t)(index >> 32)};
    uint32_t o[4];
    philox(ctr, key, o);
    int j = (int)(k & 1) * 2;
    uint64_t x = (uint64_t)o[j] << 32 | o[j + 1];
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

/* First `count` values of Stream(seed, index): kind 0 = u64, 1 = below(n),
 * 2 = uniform (as u64 bits). For the known-answer tests. */
void gen_stream_values(uint64_t seed, uint64_t index, int32_t kind, uint64_t n, int64_t count,
                       uint64_t *out) {
    Stream s;
    st_init(&s, seed, index);
    for (int64_t i = 0; i < count; i++) {
        if (kind == 0) out[i] = st_u64(&s);
        else if (kind == 1) out[i] = st_below(&s, n);
        else {
            double u = st_uniform(&s);
            memcpy(&out[i], &u, 8);
        }
    }
}

/* Stream::permutation, rng.rs:101-109 (Fisher-Yates from the top). */
void gen_permutation(uint64_t seed, uint64_t index, int64_t n, int64_t *perm) {
    Stream s;
    st_init(&s, seed, index);
    for (int64_t i = 0; i < n; i++) perm[i] = i;
    for (int64_t i = n - 1; i >= 1; i--) {
        int64_t j = (int64_t)st_below(&s, (uint64_t)i + 1);
        int64_t t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
    }
}

/* ---- helpers ------------------------------------------------------------ */

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

/* LSD radix sort of 64-bit keys on their low `bits` bits. */
static void radix_sort_u64(uint64_t *a, uint64_t n, int bits) {
    uint64_t *tmp = malloc((n ? n : 1) * sizeof *tmp);
    uint64_t *src = a, *dst = tmp;
    for (int shift = 0; shift < bits; shift += 8) {
        uint64_t cnt[257] = {0};
        for (uint64_t i = 0; i < n; i++) cnt[((src[i] >> shift) & 255) + 1]++;
        for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
        for (uint64_t i = 0; i < n; i++) dst[cnt[(src[i] >> shift) & 255]++] = src[i];
        uint64_t *t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, n * sizeof *a);
    free(tmp);
}

static int bits_for(uint64_t n) {
    int b = 0;
    while (b < 64 && (1ull << b) < n) b++;
    return b;
}

/* Sorts (from, to) pairs ascending, as EdgeList::sort, generate.rs:47-57. */
static void sort_pairs(uint32_t *from, uint32_t *to, uint64_t m, uint64_t n) {
    int b = bits_for(n ? n : 1);
    uint64_t *k = malloc((m ? m : 1) * sizeof *k);
    for (uint64_t i = 0; i < m; i++) k[i] = (uint64_t)from[i] << b | to[i];
    if (m > 4096) radix_sort_u64(k, m, 2 * b);
    else qsort(k, m, sizeof *k, cmp_u64);
    uint64_t mask = (1ull << b) - 1;
    for (uint64_t i = 0; i < m; i++) {
        from[i] = (uint32_t)(k[i] >> b);
        to[i] = (uint32_t)(k[i] & mask);
    }
    free(k);
}

/* ---- G(n, m): generate.rs:150-290 --------------------------------------- */

static uint64_t pair_count(uint64_t n, int directed, int loops) {
    if (directed) return loops ? n * n : n * (n ? n - 1 : 0);
    return loops ? n * (n + 1) / 2 : n * (n ? n - 1 : 0) / 2;
}

static uint64_t isqrt_u64(uint64_t x) {
    uint64_t r = (uint64_t)sqrt((double)x);
    while (r > 0 && (r > UINT32_MAX || r * r > x)) r--;
    while ((r + 1) <= UINT32_MAX && (r + 1) * (r + 1) <= x) r++;
    return r;
}

static void decode(uint64_t k, uint64_t n, int directed, int loops, uint64_t *a, uint64_t *b) {
    if (directed && loops) { *a = k / n; *b = k % n; return; }
    if (directed) {
        uint64_t row = k / (n - 1), col = k % (n - 1);
        *a = row;
        *b = col >= row ? col + 1 : col;
        return;
    }
    if (!loops) {
        uint64_t c = (1 + isqrt_u64(1 + 8 * k)) / 2;
        while (c * (c - 1) / 2 > k) c--;
        while ((c + 1) * c / 2 <= k) c++;
        *a = k - c * (c - 1) / 2;
        *b = c;
        return;
    }
    uint64_t c = (isqrt_u64(1 + 8 * k) - 1) / 2;
    while (c * (c + 1) / 2 > k) c--;
    while ((c + 1) * (c + 2) / 2 <= k) c++;
    *a = k - c * (c + 1) / 2;
    *b = c;
}

/* Open-addressing set of u64 (keys < UINT64_MAX). */
typedef struct { uint64_t *slot; uint64_t cap; } Set;
static void set_init(Set *s, uint64_t want) {
    uint64_t cap = 16;
    while (cap < 2 * want + 16) cap <<= 1;
    s->cap = cap;
    s->slot = malloc(cap * sizeof *s->slot);
    memset(s->slot, 0xFF, cap * sizeof *s->slot);
}
static int set_insert(Set *s, uint64_t x) {
    uint64_t h = (x * 0x9E3779B97F4A7C15ull) & (s->cap - 1);
    for (;;) {
        if (s->slot[h] == UINT64_MAX) { s->slot[h] = x; return 1; }
        if (s->slot[h] == x) return 0;
        h = (h + 1) & (s->cap - 1);
    }
}
static int set_contains(const Set *s, uint64_t x) {
    uint64_t h = (x * 0x9E3779B97F4A7C15ull) & (s->cap - 1);
    for (;;) {
        if (s->slot[h] == UINT64_MAX) return 0;
        if (s->slot[h] == x) return 1;
        h = (h + 1) & (s->cap - 1);
    }
}

/* gnm_one, generate.rs:255-290, with Stream(seed, index) (replicate r uses
 * index r, generate.rs:292-305). Writes m edges; returns m or -1. */
int64_t gen_gnm(int64_t n, int64_t m, int32_t directed, int32_t loops, uint64_t seed,
                uint64_t index, uint32_t *from, uint32_t *to) {
    uint64_t nn = (uint64_t)n, mm = (uint64_t)m;
    uint64_t total = pair_count(nn, directed, loops);
    if (mm > total) return -1;
    Stream s;
    st_init(&s, seed, index);
    uint64_t *chosen = malloc((mm ? mm : 1) * sizeof *chosen);
    uint64_t k = 0;
    if (mm * 2 <= total) {
        Set set;
        set_init(&set, mm);
        for (uint64_t j = total - mm; j < total; j++) {
            uint64_t t = st_below(&s, j + 1);
            if (!set_insert(&set, t)) set_insert(&set, j);
        }
        for (uint64_t h = 0; h < set.cap; h++)
            if (set.slot[h] != UINT64_MAX) chosen[k++] = set.slot[h];
        free(set.slot);
    } else {
        Set skip;
        set_init(&skip, total - mm);
        for (uint64_t j = mm; j < total; j++) {
            uint64_t t = st_below(&s, j + 1);
            if (!set_insert(&skip, t)) set_insert(&skip, j);
        }
        for (uint64_t x = 0; x < total; x++)
            if (!set_contains(&skip, x)) chosen[k++] = x;
        free(skip.slot);
    }
    qsort(chosen, k, sizeof *chosen, cmp_u64);
    for (uint64_t i = 0; i < k; i++) {
        uint64_t a, b;
        decode(chosen[i], nn, directed, loops, &a, &b);
        from[i] = (uint32_t)a;
        to[i] = (uint32_t)b;
    }
    free(chosen);
    sort_pairs(from, to, k, nn);
    return (int64_t)k;
}

/* ---- Watts-Strogatz: small_world_one, generate.rs:533-581 -------------- */

typedef struct { uint32_t *v; uint32_t len, cap; } NSet;
static int ns_has(const NSet *s, uint32_t x) {
    for (uint32_t i = 0; i < s->len; i++) if (s->v[i] == x) return 1;
    return 0;
}
static void ns_add(NSet *s, uint32_t x) {
    if (ns_has(s, x)) return;
    if (s->len == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 8;
        s->v = realloc(s->v, s->cap * sizeof *s->v);
    }
    s->v[s->len++] = x;
}
static void ns_del(NSet *s, uint32_t x) {
    for (uint32_t i = 0; i < s->len; i++)
        if (s->v[i] == x) { s->v[i] = s->v[--s->len]; return; }
}

/* Writes n*k/2 edges (u < v, sorted); returns the count or -1. */
int64_t gen_small_world(int64_t n, int64_t k, double p, uint64_t seed, uint64_t index,
                        uint32_t *from, uint32_t *to) {
    if (n <= 0 || k % 2 == 1 || k >= n) return -1;
    Stream s;
    st_init(&s, seed, index);
    NSet *adj = calloc((size_t)n, sizeof *adj);
    int64_t half = k / 2;
    for (int64_t u = 0; u < n; u++)
        for (int64_t j = 1; j <= half; j++) {
            int64_t v = (u + j) % n;
            ns_add(&adj[u], (uint32_t)v);
            ns_add(&adj[v], (uint32_t)u);
        }
    for (int64_t j = 1; j <= half; j++) {
        for (int64_t u = 0; u < n; u++) {
            uint32_t v = (uint32_t)((u + j) % n);
            if (st_uniform(&s) >= p) continue;
            if ((int64_t)adj[u].len >= n - 1) continue;
            if (!ns_has(&adj[u], v)) continue;
            uint32_t w;
            for (;;) {
                w = (uint32_t)st_below(&s, (uint64_t)n);
                if ((int64_t)w != u && !ns_has(&adj[u], w)) break;
            }
            ns_del(&adj[u], v);
            ns_del(&adj[v], (uint32_t)u);
            ns_add(&adj[u], w);
            ns_add(&adj[w], (uint32_t)u);
        }
    }
    int64_t m = 0;
    for (int64_t u = 0; u < n; u++) {
        for (uint32_t i = 0; i < adj[u].len; i++) {
            if ((int64_t)adj[u].v[i] > u) {
                from[m] = (uint32_t)u;
                to[m] = adj[u].v[i];
                m++;
            }
        }
        free(adj[u].v);
    }
    free(adj);
    sort_pairs(from, to, (uint64_t)m, (uint64_t)n);
    return m;
}

/* ---- Barabasi-Albert: preferential_one, generate.rs:602-631 ------------ */

/* Writes m*(n-m) edges; returns the count or -1. */
int64_t gen_preferential(int64_t n, int64_t m, uint64_t seed, uint64_t index, uint32_t *from,
                         uint32_t *to) {
    if (m < 1 || n <= m) return -1;
    Stream s;
    st_init(&s, seed, index);
    uint64_t cnt = 0;
    uint32_t *pool = malloc((size_t)(2 * m * n) * sizeof *pool);
    uint64_t plen = 0;
    for (int64_t leaf = 1; leaf <= m; leaf++) {
        from[cnt] = 0;
        to[cnt] = (uint32_t)leaf;
        cnt++;
        pool[plen++] = 0;
        pool[plen++] = (uint32_t)leaf;
    }
    uint32_t *targets = malloc((size_t)m * sizeof *targets);
    for (int64_t source = m + 1; source < n; source++) {
        int64_t nt = 0;
        while (nt < m) {
            uint32_t t = pool[st_below(&s, plen)];
            int dup = 0;
            for (int64_t i = 0; i < nt; i++) if (targets[i] == t) { dup = 1; break; }
            if (!dup) targets[nt++] = t;
        }
        for (int64_t i = 0; i < m; i++) {
            from[cnt] = targets[i];
            to[cnt] = (uint32_t)source;
            cnt++;
            pool[plen++] = targets[i];
            pool[plen++] = (uint32_t)source;
        }
    }
    free(targets);
    free(pool);
    sort_pairs(from, to, cnt, (uint64_t)n);
    return (int64_t)cnt;
}

/* ---- R-MAT (this library's definition) ----------------------------------
 * Edge e of 2^scale * edge_factor draws `scale` uniforms from
 * Stream(seed, 0), uniforms e*scale .. e*scale+scale-1 in order (so a
 * single sequential stream reproduces it). Draw i picks the quadrant for
 * bit scale-1-i of (row, col): u < a -> (0,0); u < a+b -> (0,1);
 * u < a+b+c -> (1,0); else (1,1). Row is the edge source, col its target.
 * No noise and no node permutation. Self-loops are dropped; directed
 * graphs keep distinct (row, col) pairs, undirected graphs distinct
 * (min, max) pairs stored with the smaller node first; the result is
 * sorted. Nodes without edges stay in the graph. */

typedef struct {
    uint32_t key[2];
    int scale;
    double a, ab, abc;
    int directed;
    uint64_t lo, hi;
    uint64_t *keys;
    uint64_t kept;
} RmatJob;

static void *rmat_worker(void *arg) {
    RmatJob *j = arg;
    uint64_t kept = 0;
    int sc = j->scale;
    for (uint64_t e = j->lo; e < j->hi; e++) {
        uint64_t row = 0, col = 0;
        uint64_t base = e * (uint64_t)sc;
        for (int i = 0; i < sc; i++) {
            double u = uniform_at(j->key, 0, base + (uint64_t)i);
            uint64_t bit = 1ull << (sc - 1 - i);
            if (u < j->a) {
            } else if (u < j->ab) {
                col |= bit;
            } else if (u < j->abc) {
                row |= bit;
            } else {
                row |= bit;
                col |= bit;
            }
        }
        if (row == col) continue;
        if (!j->directed && row > col) { uint64_t t = row; row = col; col = t; }
        j->keys[j->lo + kept++] = row << sc | col;
    }
    j->kept = kept;
    return NULL;
}

/* Returns the edge count and heap arrays in *from / *to (free with
 * gen_free), or -1. */
int64_t gen_rmat(int32_t scale, int64_t edge_factor, double a, double b, double c,
                 int32_t directed, uint64_t seed, int32_t threads, uint32_t **from,
                 uint32_t **to) {
    if (scale < 1 || scale > 31 || edge_factor < 1) return -1;
    uint64_t total = ((uint64_t)1 << scale) * (uint64_t)edge_factor;
    uint64_t *keys = malloc(total * sizeof *keys);
    if (!keys) return -1;
    if (threads < 1) threads = 1;
    RmatJob *jobs = calloc((size_t)threads, sizeof *jobs);
    pthread_t *tid = calloc((size_t)threads, sizeof *tid);
    for (int t = 0; t < threads; t++) {
        RmatJob *j = &jobs[t];
        j->key[0] = (uint32_t)seed;
        j->key[1] = (uint32_t)(seed >> 32);
        j->scale = scale;
        j->a = a;
        j->ab = a + b;
        j->abc = a + b + c;
        j->directed = directed;
        j->lo = total * (uint64_t)t / (uint64_t)threads;
        j->hi = total * (uint64_t)(t + 1) / (uint64_t)threads;
        j->keys = keys;
        pthread_create(&tid[t], NULL, rmat_worker, j);
    }
    uint64_t m = 0;
    for (int t = 0; t < threads; t++) {
        pthread_join(tid[t], NULL);
        memmove(keys + m, keys + jobs[t].lo, jobs[t].kept * sizeof *keys);
        m += jobs[t].kept;
    }
    free(jobs);
    free(tid);
    radix_sort_u64(keys, m, 2 * scale);
    uint64_t u = 0;
    for (uint64_t i = 0; i < m; i++)
        if (u == 0 || keys[i] != keys[u - 1]) keys[u++] = keys[i];
    *from = malloc((u ? u : 1) * sizeof **from);
    *to = malloc((u ? u : 1) * sizeof **to);
    uint64_t mask = ((uint64_t)1 << scale) - 1;
    for (uint64_t i = 0; i < u; i++) {
        (*from)[i] = (uint32_t)(keys[i] >> scale);
        (*to)[i] = (uint32_t)(keys[i] & mask);
    }
    free(keys);
    return (int64_t)u;
}

void gen_free(void *p) { free(p); }

/* FNV-1a 64 over (from, to) pairs as little-endian u32s: a fingerprint
 * for committed edge lists. */
uint64_t gen_fingerprint(const uint32_t *from, const uint32_t *to, int64_t m) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (int64_t i = 0; i < m; i++) {
        uint32_t x[2] = {from[i], to[i]};
        const unsigned char *p = (const unsigned char *)x;
        for (int k = 0; k < 8; k++) {
            h ^= p[k];
            h *= 0x100000001b3ull;
        }
    }
    return h;
}
