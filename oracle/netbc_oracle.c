/*
 * CPU oracle for knappnet's exact Brandes betweenness (net_betweenness).
 *
 * TEST INFRASTRUCTURE ONLY. This file restates, in plain C11, the
 * reference engine's CPU algorithm so that the device library can be
 * checked against it. Only the tests, the smoke check and the benchmark's
 * CPU-baseline leg load it; the product library never links or calls it.
 *
 * Every function cites the reference file:line it follows (paths are
 * relative to the knappnet source tree). Floating-point operations are
 * written in the reference's order and the file is compiled with
 * -ffp-contract=off and without fast-math, so that FP64 results follow the
 * reference bit for bit on the same inputs.
 *
 * Pinning (what checks each part; see tests/):
 *   graph build      graph.rs:350-412 known answers, host/device CSR equality
 *   path lengths     weights.rs:290-321 known answers
 *   BFS/Dial/Dijk.   betweenness.rs:387-517 known answers, the 28 committed
 *                    definition/NetworkX goldens, brute-force exact-rational
 *                    definition on random small graphs, the sum identity
 *   sum_in_order     parallel.rs:653-665 thread-count invariance
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define ORC_OK 0
#define ORC_FAILED 1
#define U32_MAX_ 0xFFFFFFFFu

/* ---- errors: lib.rs:60-71 (write_err), error.rs:62-99 (texts) ---------- */

static void write_err(char *err, size_t err_len, const char *msg) {
    if (!err || err_len == 0) return;
    size_t n = strlen(msg);
    if (n > err_len - 1) n = err_len - 1;
    memcpy(err, msg, n);
    err[n] = 0;
}

/* Rust's `Display` for f64: the shortest digits that round-trip, written
 * positionally (no exponent). Used in the NonPositiveWeight message,
 * error.rs:81-85. */
static void fmt_f64(double x, char *out, size_t cap) {
    if (isnan(x)) { snprintf(out, cap, "NaN"); return; }
    if (isinf(x)) { snprintf(out, cap, x < 0 ? "-inf" : "inf"); return; }
    char sign[2] = {0, 0};
    if (signbit(x)) { sign[0] = '-'; x = -x; }
    if (x == 0.0) { snprintf(out, cap, "%s0", sign); return; }
    char buf[64];
    int p;
    for (p = 1; p <= 17; p++) {
        snprintf(buf, sizeof buf, "%.*e", p - 1, x);
        if (strtod(buf, NULL) == x) break;
    }
    /* buf = d[.ddd]e[+-]XX */
    char digits[32];
    int nd = 0;
    char *e = strchr(buf, 'e');
    for (char *c = buf; c < e; c++)
        if (*c >= '0' && *c <= '9') digits[nd++] = *c;
    int exp10 = atoi(e + 1);
    while (nd > 1 && digits[nd - 1] == '0') nd--;
    char res[400];
    int r = 0;
    if (exp10 >= 0) {
        for (int i = 0; i <= exp10; i++) res[r++] = i < nd ? digits[i] : '0';
        if (nd > exp10 + 1) {
            res[r++] = '.';
            for (int i = exp10 + 1; i < nd; i++) res[r++] = digits[i];
        }
    } else {
        res[r++] = '0';
        res[r++] = '.';
        for (int i = 0; i < -exp10 - 1; i++) res[r++] = '0';
        for (int i = 0; i < nd; i++) res[r++] = digits[i];
    }
    res[r] = 0;
    snprintf(out, cap, "%s%s", sign, res);
}

/* ---- graph: graph.rs:18-234 -------------------------------------------- */

typedef struct {
    uint64_t *offsets; /* n + 1 */
    uint32_t *targets; /* arcs */
    double *weights;   /* arcs or NULL */
    uint64_t arcs;
} OCsr;

typedef struct OGraph {
    uint64_t n, m;
    int directed;
    uint32_t *src, *tgt; /* input order */
    double *w;           /* input order or NULL */
    OCsr out, inc;       /* inc only when directed */
} OGraph;

/* Csr::build, graph.rs:72-118: two stable counting sorts, first by target
 * then by source, so each row ends up sorted by target with parallel arcs
 * in input order. */
static void csr_build(uint64_t n, const uint32_t *from, const uint32_t *to,
                      const double *weights, uint64_t m, OCsr *c) {
    uint64_t *by_target = calloc(n + 1, sizeof *by_target);
    for (uint64_t a = 0; a < m; a++) by_target[(uint64_t)to[a] + 1]++;
    for (uint64_t i = 0; i < n; i++) by_target[i + 1] += by_target[i];
    uint64_t *order = malloc((m ? m : 1) * sizeof *order);
    uint64_t *cursor = malloc((n + 1) * sizeof *cursor);
    memcpy(cursor, by_target, (n + 1) * sizeof *cursor);
    for (uint64_t a = 0; a < m; a++) order[cursor[to[a]]++] = a;

    c->offsets = calloc(n + 1, sizeof *c->offsets);
    for (uint64_t a = 0; a < m; a++) c->offsets[(uint64_t)from[a] + 1]++;
    for (uint64_t i = 0; i < n; i++) c->offsets[i + 1] += c->offsets[i];
    memcpy(cursor, c->offsets, (n + 1) * sizeof *cursor);
    c->targets = malloc((m ? m : 1) * sizeof *c->targets);
    c->weights = weights ? malloc((m ? m : 1) * sizeof *c->weights) : NULL;
    for (uint64_t i = 0; i < m; i++) {
        uint64_t a = order[i];
        uint64_t slot = cursor[from[a]]++;
        c->targets[slot] = to[a];
        if (weights) c->weights[slot] = weights[a];
    }
    c->arcs = m;
    free(by_target);
    free(order);
    free(cursor);
}

static void csr_free(OCsr *c) {
    free(c->offsets);
    free(c->targets);
    free(c->weights);
    memset(c, 0, sizeof *c);
}

void orc_graph_free(OGraph *g) {
    if (!g) return;
    csr_free(&g->out);
    csr_free(&g->inc);
    free(g->src);
    free(g->tgt);
    free(g->w);
    free(g);
}

/* net_graph_new, lib.rs:133-164, and Graph::from_edges, graph.rs:142-234,
 * with the same order of checks and the same messages. */
int32_t orc_graph_new(int64_t n, const int64_t *src, const int64_t *tgt,
                      const double *weights, int64_t m, int32_t directed,
                      OGraph **out, char *err, size_t err_len) {
    char msg[512];
    if (n < 0 || m < 0) {
        write_err(err, err_len, "negative node or edge count");
        return ORC_FAILED;
    }
    if ((uint64_t)n > (uint64_t)U32_MAX_) {
        snprintf(msg, sizeof msg, "node count (%lld) exceeds the supported maximum", (long long)n);
        write_err(err, err_len, msg);
        return ORC_FAILED;
    }
    if (weights) {
        for (int64_t e = 0; e < m; e++) {
            if (!isfinite(weights[e])) {
                snprintf(msg, sizeof msg, "edge %lld has a missing, NaN, or infinite weight",
                         (long long)e);
                write_err(err, err_len, msg);
                return ORC_FAILED;
            }
        }
    }
    if ((uint64_t)m > (uint64_t)(U32_MAX_ / 2)) {
        snprintf(msg, sizeof msg, "edge count (%lld) exceeds the supported maximum", (long long)m);
        write_err(err, err_len, msg);
        return ORC_FAILED;
    }
    OGraph *g = calloc(1, sizeof *g);
    g->n = (uint64_t)n;
    g->m = (uint64_t)m;
    g->directed = directed != 0;
    g->src = malloc((m ? m : 1) * sizeof *g->src);
    g->tgt = malloc((m ? m : 1) * sizeof *g->tgt);
    for (int64_t e = 0; e < m; e++) {
        int64_t ends[2] = {src[e], tgt[e]};
        for (int k = 0; k < 2; k++) {
            if (ends[k] < 0 || (uint64_t)ends[k] >= (uint64_t)n) {
                snprintf(msg, sizeof msg,
                         "edge %lld refers to node index %lld, but valid indices are 0..%lld",
                         (long long)e, (long long)ends[k], (long long)n);
                write_err(err, err_len, msg);
                orc_graph_free(g);
                return ORC_FAILED;
            }
        }
        g->src[e] = (uint32_t)src[e];
        g->tgt[e] = (uint32_t)tgt[e];
    }
    if (weights) {
        g->w = malloc((m ? m : 1) * sizeof *g->w);
        memcpy(g->w, weights, (size_t)m * sizeof *g->w);
    }
    if (g->directed) {
        csr_build(g->n, g->src, g->tgt, g->w, g->m, &g->out);
        csr_build(g->n, g->tgt, g->src, g->w, g->m, &g->inc);
    } else {
        /* Each non-loop edge becomes two arcs; a loop stays one arc
         * (graph.rs:201-223). */
        uint64_t loops = 0;
        for (uint64_t e = 0; e < g->m; e++) loops += g->src[e] == g->tgt[e];
        uint64_t arcs = 2 * g->m - loops;
        uint32_t *from = malloc((arcs ? arcs : 1) * sizeof *from);
        uint32_t *to = malloc((arcs ? arcs : 1) * sizeof *to);
        double *w = g->w ? malloc((arcs ? arcs : 1) * sizeof *w) : NULL;
        uint64_t k = 0;
        for (uint64_t e = 0; e < g->m; e++) {
            from[k] = g->src[e];
            to[k] = g->tgt[e];
            if (w) w[k] = g->w[e];
            k++;
            if (g->src[e] != g->tgt[e]) {
                from[k] = g->tgt[e];
                to[k] = g->src[e];
                if (w) w[k] = g->w[e];
                k++;
            }
        }
        csr_build(g->n, from, to, w, arcs, &g->out);
        free(from);
        free(to);
        free(w);
    }
    *out = g;
    return ORC_OK;
}

int64_t orc_graph_node_count(const OGraph *g) { return g ? (int64_t)g->n : -1; }
int64_t orc_graph_edge_count(const OGraph *g) { return g ? (int64_t)g->m : -1; }
int64_t orc_graph_arc_count(const OGraph *g, int32_t incoming) {
    if (!g) return -1;
    return (int64_t)((incoming && g->directed) ? g->inc.arcs : g->out.arcs);
}

/* Copies one CSR direction out (incoming == 0: out-arcs; otherwise the
 * in-arcs, which for undirected graphs are the out-arcs, graph.rs:281-285). */
void orc_graph_csr(const OGraph *g, int32_t incoming, uint64_t *offsets, uint32_t *targets,
                   double *weights) {
    const OCsr *c = (incoming && g->directed) ? &g->inc : &g->out;
    memcpy(offsets, c->offsets, (g->n + 1) * sizeof *offsets);
    memcpy(targets, c->targets, c->arcs * sizeof *targets);
    if (weights && c->weights) memcpy(weights, c->weights, c->arcs * sizeof *weights);
}

/* ---- path lengths: weights.rs:43-51, 89-131, 151-153 ------------------- */

enum { LEN_UNIFORM = 0, LEN_INTEGER = 1, LEN_REAL = 2 };
#define DIAL_MAX 1024u

typedef struct {
    int kind;
    uint32_t max;
    uint32_t *ilens; /* arcs, CSR order (Integer) */
    double *rlens;   /* arcs, CSR order (Real) */
} Lengths;

static void lengths_free(Lengths *L) {
    free(L->ilens);
    free(L->rlens);
}

/* algorithm: 0 = Auto, 1 = Bfs, 2 = Dial, 3 = Dijkstra (weights.rs:54-66). */
static int path_lengths(const OGraph *g, int weighted, int kind, int algorithm, Lengths *L,
                        char *err, size_t err_len) {
    memset(L, 0, sizeof *L);
    if (!weighted || algorithm == 1) { L->kind = LEN_UNIFORM; return ORC_OK; }
    if (!g->w) { write_err(err, err_len, "the graph has no edge weights"); return ORC_FAILED; }
    for (uint64_t e = 0; e < g->m; e++) {
        if (g->w[e] <= 0.0) {
            char num[400], msg[600];
            fmt_f64(g->w[e], num, sizeof num);
            snprintf(msg, sizeof msg,
                     "edge %llu has weight %s; this algorithm treats weights as distances and "
                     "requires them to be strictly positive",
                     (unsigned long long)e, num);
            write_err(err, err_len, msg);
            return ORC_FAILED;
        }
    }
    uint64_t A = g->out.arcs;
    double *lens = malloc((A ? A : 1) * sizeof *lens);
    for (uint64_t a = 0; a < A; a++)
        lens[a] = kind == 1 ? 1.0 / g->out.weights[a] : g->out.weights[a];
    int uniform = 1;
    for (uint64_t a = 1; a < A; a++)
        if (!(lens[a - 1] == lens[a])) { uniform = 0; break; }
    int is_int = 1;
    uint64_t imax = 0;
    for (uint64_t a = 0; a < A; a++) {
        double x = lens[a];
        if (x - trunc(x) == 0.0 && x >= 1.0 && x <= 4294967295.0) {
            if ((uint64_t)x > imax) imax = (uint64_t)x;
        } else {
            is_int = 0;
            break;
        }
    }
    int use_int = 0;
    if (algorithm == 0) {
        if (uniform) { free(lens); L->kind = LEN_UNIFORM; return ORC_OK; }
        use_int = is_int && imax <= DIAL_MAX;
    } else if (algorithm == 2) {
        if (!is_int) {
            free(lens);
            write_err(err, err_len, "Dial's algorithm needs positive integer path lengths");
            return ORC_FAILED;
        }
        use_int = 1;
    }
    if (use_int) {
        L->kind = LEN_INTEGER;
        L->max = (uint32_t)imax;
        L->ilens = malloc((A ? A : 1) * sizeof *L->ilens);
        for (uint64_t a = 0; a < A; a++) L->ilens[a] = (uint32_t)lens[a];
        free(lens);
    } else {
        L->kind = LEN_REAL;
        L->rlens = lens;
    }
    return ORC_OK;
}

int32_t orc_path_lengths(const OGraph *g, int32_t weighted, int32_t weights_as, int32_t algorithm,
                         int32_t *kind, uint32_t *max, char *err, size_t err_len) {
    Lengths L;
    if (path_lengths(g, weighted, weights_as, algorithm, &L, err, err_len)) return ORC_FAILED;
    *kind = L.kind;
    *max = L.max;
    lengths_free(&L);
    return ORC_OK;
}

/* ---- indexed 4-ary heap: weights.rs:155-277 ---------------------------- */

typedef struct { double d; uint32_t v; } HItem;
typedef struct { HItem *items; uint64_t len; uint32_t *pos; } Heap;

static int h_less(HItem a, HItem b) {
    if (a.d < b.d) return 1;
    if (a.d > b.d) return 0;
    return a.v < b.v;
}

static void h_sift_up(Heap *h, uint64_t i) {
    HItem item = h->items[i];
    while (i > 0) {
        uint64_t parent = (i - 1) / 4;
        HItem p = h->items[parent];
        if (!h_less(item, p)) break;
        h->items[i] = p;
        h->pos[p.v] = (uint32_t)i;
        i = parent;
    }
    h->items[i] = item;
    h->pos[item.v] = (uint32_t)i;
}

static void h_sift_down(Heap *h, uint64_t i) {
    uint64_t len = h->len;
    HItem item = h->items[i];
    for (;;) {
        uint64_t first = i * 4 + 1;
        if (first >= len) break;
        uint64_t best = first;
        uint64_t end = first + 4 < len ? first + 4 : len;
        for (uint64_t c = first + 1; c < end; c++)
            if (h_less(h->items[c], h->items[best])) best = c;
        HItem b = h->items[best];
        if (!h_less(b, item)) break;
        h->items[i] = b;
        h->pos[b.v] = (uint32_t)i;
        i = best;
    }
    h->items[i] = item;
    h->pos[item.v] = (uint32_t)i;
}

static void h_push_or_decrease(Heap *h, uint32_t node, double dist) {
    uint32_t p = h->pos[node];
    uint64_t i;
    if (p == U32_MAX_) {
        h->items[h->len] = (HItem){dist, node};
        i = h->len++;
    } else {
        i = p;
        HItem cand = {dist, node};
        if (!h_less(cand, h->items[i])) return;
        h->items[i].d = dist;
    }
    h_sift_up(h, i);
}

static int h_pop(Heap *h, HItem *top) {
    if (h->len == 0) return 0;
    *top = h->items[0];
    h->pos[top->v] = U32_MAX_;
    HItem last = h->items[--h->len];
    if (h->len > 0) {
        h->items[0] = last;
        h_sift_down(h, 0);
    }
    return 1;
}

/* ---- per-thread workspace: betweenness.rs:122-159 ---------------------- */

typedef struct { uint32_t *a; uint64_t len, cap; } U32Vec;
static void vpush(U32Vec *v, uint32_t x) {
    if (v->len == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 16;
        v->a = realloc(v->a, v->cap * sizeof *v->a);
    }
    v->a[v->len++] = x;
}

typedef struct {
    uint64_t n;
    double *bc; /* the accumulator lent for the current source */
    double *sigma, *coef, *dist;
    uint32_t *hops, *rank, *order;
    uint64_t norder;
    uint64_t *idist;
    U32Vec touched;
    U32Vec *buckets;
    uint32_t nbuckets;
    Heap heap;
    /* optional per-source delta capture (debug entry) */
    double *delta_out;
    /* edge betweenness (reach.rs:401-448): coef holds delta, eb is the
     * lent per-edge accumulator, arc_edge maps CSR arcs to input edges */
    int edge;
    double *eb;
    const uint32_t *arc_edge;
} Workspace;

static void ws_init(Workspace *ws, uint64_t n) {
    memset(ws, 0, sizeof *ws);
    ws->n = n;
    uint64_t nn = n ? n : 1;
    ws->sigma = calloc(nn, sizeof(double));
    ws->coef = calloc(nn, sizeof(double));
    ws->dist = malloc(nn * sizeof(double));
    ws->hops = malloc(nn * sizeof(uint32_t));
    ws->rank = malloc(nn * sizeof(uint32_t));
    ws->order = malloc(nn * sizeof(uint32_t));
    ws->idist = malloc(nn * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) {
        ws->dist[i] = INFINITY;
        ws->hops[i] = U32_MAX_;
        ws->rank[i] = U32_MAX_;
        ws->idist[i] = UINT64_MAX;
    }
    ws->heap.items = malloc(nn * sizeof(HItem));
    ws->heap.pos = malloc(nn * sizeof(uint32_t));
    for (uint64_t i = 0; i < n; i++) ws->heap.pos[i] = U32_MAX_;
}

static void ws_free(Workspace *ws) {
    free(ws->sigma); free(ws->coef); free(ws->dist); free(ws->hops);
    free(ws->rank); free(ws->order); free(ws->idist); free(ws->touched.a);
    for (uint32_t i = 0; i < ws->nbuckets; i++) free(ws->buckets[i].a);
    free(ws->buckets);
    free(ws->heap.items); free(ws->heap.pos);
}

/* Adds the dependency of v to the lent accumulator, or records it for the
 * debug entry. */
static inline void credit(Workspace *ws, uint32_t v, uint64_t s, double delta) {
    if (ws->delta_out) ws->delta_out[v] = delta;
    if (v != s && ws->bc) ws->bc[v] += delta;
}

/* approx_eq with EPSILON = 1e-10, betweenness.rs:44, 117-120 (and
 * reach.rs:85, 255-257). */
static inline int approx_eq(double a, double b) {
    double m = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
    return fabs(a - b) <= 1e-10 * m;
}

/* EdgeWorkspace::accumulate, reach.rs:401-448: farthest nodes first, every
 * arc (v, w) on a shortest path receives c = sigma_v / sigma_w * (1 +
 * delta_w), credited to its input edge; delta_v is the sum of its arcs' c
 * in CSR order. kind: 0 hops, 1 integer distances, 2 real distances. */
static void backward_edges(Workspace *ws, const OGraph *g, int kind, const uint32_t *ilens,
                           const double *rlens) {
    const OCsr *adj = &g->out;
    for (uint64_t i = ws->norder; i-- > 0;) {
        uint32_t v = ws->order[i];
        uint32_t rank_v = ws->rank[v];
        double dv = 0.0;
        for (uint64_t k = adj->offsets[v]; k < adj->offsets[v + 1]; k++) {
            uint32_t w = adj->targets[k];
            int on_path;
            if (kind == 0)
                on_path = ws->hops[w] != U32_MAX_ && ws->hops[w] == ws->hops[v] + 1;
            else if (kind == 1)
                on_path = ws->rank[w] != U32_MAX_ && ws->rank[w] > rank_v &&
                          ws->idist[v] + (uint64_t)ilens[k] == ws->idist[w];
            else
                on_path = ws->rank[w] != U32_MAX_ && ws->rank[w] > rank_v &&
                          approx_eq(ws->dist[v] + rlens[k], ws->dist[w]);
            if (on_path) {
                double c = ws->sigma[v] / ws->sigma[w] * (1.0 + ws->coef[w]);
                if (ws->eb) ws->eb[ws->arc_edge[k]] += c;
                dv += c;
            }
        }
        ws->coef[v] = dv;
        if (ws->delta_out) ws->delta_out[v] = dv;
    }
}

/* Workspace::single_source_bfs, betweenness.rs:161-210. */
static void single_source_bfs(Workspace *ws, const OGraph *g, uint64_t s) {
    const OCsr *adj = &g->out;
    ws->norder = 0;
    ws->hops[s] = 0;
    ws->sigma[s] = 1.0;
    ws->order[ws->norder++] = (uint32_t)s;
    uint64_t head = 0;
    while (head < ws->norder) {
        uint32_t v = ws->order[head++];
        uint32_t next = ws->hops[v] + 1;
        for (uint64_t a = adj->offsets[v]; a < adj->offsets[v + 1]; a++) {
            uint32_t w = adj->targets[a];
            if (ws->hops[w] == U32_MAX_) {
                ws->hops[w] = next;
                ws->order[ws->norder++] = w;
            }
            if (ws->hops[w] == next) ws->sigma[w] += ws->sigma[v];
        }
    }
    if (ws->edge) {
        backward_edges(ws, g, 0, NULL, NULL);
        return;
    }
    for (uint64_t i = ws->norder; i-- > 0;) {
        uint32_t v = ws->order[i];
        uint32_t next = ws->hops[v] + 1;
        double dv = 0.0;
        for (uint64_t a = adj->offsets[v]; a < adj->offsets[v + 1]; a++) {
            uint32_t w = adj->targets[a];
            if (ws->hops[w] == next) dv += ws->coef[w];
        }
        double delta = ws->sigma[v] * dv;
        ws->coef[v] = (1.0 + delta) / ws->sigma[v];
        credit(ws, v, s, delta);
    }
}

static void reset_bfs(Workspace *ws) {
    for (uint64_t i = 0; i < ws->norder; i++) {
        uint32_t v = ws->order[i];
        ws->hops[v] = U32_MAX_;
        ws->sigma[v] = 0.0;
        ws->coef[v] = 0.0;
    }
}

/* Workspace::single_source_dial, betweenness.rs:214-295: Dial's circular
 * buckets of width max + 1, exact integer distances, LIFO buckets. */
static void single_source_dial(Workspace *ws, const OGraph *g, const uint32_t *lens,
                               uint32_t max, uint64_t s) {
    const OCsr *adj = &g->out;
    uint64_t width = (uint64_t)max + 1;
    if (ws->nbuckets != width) {
        for (uint32_t i = 0; i < ws->nbuckets; i++) free(ws->buckets[i].a);
        free(ws->buckets);
        ws->buckets = calloc(width, sizeof(U32Vec));
        ws->nbuckets = (uint32_t)width;
    }
    ws->norder = 0;
    ws->idist[s] = 0;
    ws->sigma[s] = 1.0;
    vpush(&ws->buckets[0], (uint32_t)s);
    uint64_t pending = 1;
    ws->touched.len = 0;
    vpush(&ws->touched, (uint32_t)s);
    uint64_t cur = 0;
    while (pending > 0) {
        U32Vec *b = &ws->buckets[cur % width];
        if (b->len == 0) { cur++; continue; }
        uint32_t v = b->a[--b->len];
        pending--;
        if (ws->rank[v] != U32_MAX_ || ws->idist[v] != cur) continue;
        ws->rank[v] = (uint32_t)ws->norder;
        ws->order[ws->norder++] = v;
        for (uint64_t k = adj->offsets[v]; k < adj->offsets[v + 1]; k++) {
            uint32_t w = adj->targets[k];
            if (ws->rank[w] != U32_MAX_) continue;
            uint64_t alt = cur + (uint64_t)lens[k];
            uint64_t d = ws->idist[w];
            if (d == UINT64_MAX || alt < d) {
                if (d == UINT64_MAX) vpush(&ws->touched, w);
                ws->idist[w] = alt;
                ws->sigma[w] = ws->sigma[v];
                vpush(&ws->buckets[alt % width], w);
                pending++;
            } else if (alt == d) {
                ws->sigma[w] += ws->sigma[v];
            }
        }
    }
    if (ws->edge) {
        backward_edges(ws, g, 1, lens, NULL);
        return;
    }
    for (uint64_t i = ws->norder; i-- > 0;) {
        uint32_t v = ws->order[i];
        uint32_t rank_v = ws->rank[v];
        uint64_t dist_v = ws->idist[v];
        double dv = 0.0;
        for (uint64_t k = adj->offsets[v]; k < adj->offsets[v + 1]; k++) {
            uint32_t w = adj->targets[k];
            if (ws->rank[w] != U32_MAX_ && ws->rank[w] > rank_v &&
                dist_v + (uint64_t)lens[k] == ws->idist[w])
                dv += ws->coef[w];
        }
        double delta = ws->sigma[v] * dv;
        ws->coef[v] = (1.0 + delta) / ws->sigma[v];
        credit(ws, v, s, delta);
    }
}

static void reset_touched_int(Workspace *ws) {
    for (uint64_t i = 0; i < ws->touched.len; i++) {
        uint32_t v = ws->touched.a[i];
        ws->idist[v] = UINT64_MAX;
        ws->sigma[v] = 0.0;
        ws->coef[v] = 0.0;
        ws->rank[v] = U32_MAX_;
    }
}

/* Workspace::single_source_dijkstra, betweenness.rs:297-368. */
static void single_source_dijkstra(Workspace *ws, const OGraph *g, const double *lens,
                                   uint64_t s) {
    const OCsr *adj = &g->out;
    ws->norder = 0;
    ws->dist[s] = 0.0;
    ws->sigma[s] = 1.0;
    h_push_or_decrease(&ws->heap, (uint32_t)s, 0.0);
    ws->touched.len = 0;
    vpush(&ws->touched, (uint32_t)s);
    HItem top;
    while (h_pop(&ws->heap, &top)) {
        uint32_t v = top.v;
        double d = top.d;
        ws->rank[v] = (uint32_t)ws->norder;
        ws->order[ws->norder++] = v;
        for (uint64_t k = adj->offsets[v]; k < adj->offsets[v + 1]; k++) {
            uint32_t w = adj->targets[k];
            if (ws->rank[w] != U32_MAX_) continue;
            double alt = d + lens[k];
            double cur = ws->dist[w];
            if (isinf(cur)) {
                vpush(&ws->touched, w);
                ws->dist[w] = alt;
                ws->sigma[w] = ws->sigma[v];
                h_push_or_decrease(&ws->heap, w, alt);
            } else if (approx_eq(alt, cur)) {
                ws->sigma[w] += ws->sigma[v];
            } else if (alt < cur) {
                ws->dist[w] = alt;
                ws->sigma[w] = ws->sigma[v];
                h_push_or_decrease(&ws->heap, w, alt);
            }
        }
    }
    if (ws->edge) {
        /* reach.rs:323-370 pops a lazy binary heap ordered by (distance,
         * node); the indexed heap settles in the same order */
        backward_edges(ws, g, 2, NULL, lens);
        return;
    }
    for (uint64_t i = ws->norder; i-- > 0;) {
        uint32_t v = ws->order[i];
        uint32_t rank_v = ws->rank[v];
        double dist_v = ws->dist[v];
        double dv = 0.0;
        for (uint64_t k = adj->offsets[v]; k < adj->offsets[v + 1]; k++) {
            uint32_t w = adj->targets[k];
            if (ws->rank[w] > rank_v && ws->rank[w] != U32_MAX_ &&
                approx_eq(dist_v + lens[k], ws->dist[w]))
                dv += ws->coef[w];
        }
        double delta = ws->sigma[v] * dv;
        ws->coef[v] = (1.0 + delta) / ws->sigma[v];
        credit(ws, v, s, delta);
    }
}

static void reset_touched_real(Workspace *ws) {
    for (uint64_t i = 0; i < ws->touched.len; i++) {
        uint32_t v = ws->touched.a[i];
        ws->dist[v] = INFINITY;
        ws->sigma[v] = 0.0;
        ws->coef[v] = 0.0;
        ws->rank[v] = U32_MAX_;
    }
}

static void run_source(Workspace *ws, const OGraph *g, const Lengths *L, uint64_t s) {
    switch (L->kind) {
    case LEN_UNIFORM:
        single_source_bfs(ws, g, s);
        reset_bfs(ws);
        break;
    case LEN_INTEGER:
        single_source_dial(ws, g, L->ilens, L->max, s);
        reset_touched_int(ws);
        break;
    default:
        single_source_dijkstra(ws, g, L->rlens, s);
        reset_touched_real(ws);
        break;
    }
}

/* ---- threads: parallel.rs:28-60 ---------------------------------------- */

static uint64_t available_cores(void) {
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof set, &set) == 0) {
        int c = CPU_COUNT(&set);
        if (c > 0) return (uint64_t)c;
    }
    long c = sysconf(_SC_NPROCESSORS_ONLN);
    return c > 0 ? (uint64_t)c : 1;
}

/* resolve_threads, parallel.rs:29-40. */
static uint64_t resolve_threads(uint64_t requested, uint64_t items) {
    uint64_t t = requested == 0 ? available_cores() : requested;
    uint64_t hi = items > 1 ? items : 1;
    if (t < 1) t = 1;
    if (t > hi) t = hi;
    return t;
}

#define GRAIN (1ull << 17)
/* threads_for, parallel.rs:51-54; search_work, parallel.rs:58-60. */
static uint64_t threads_for(uint64_t requested, uint64_t items, uint64_t work) {
    uint64_t by_work = work / GRAIN;
    if (by_work < 1) by_work = 1;
    uint64_t t = resolve_threads(requested, items);
    return t < by_work ? t : by_work;
}

static uint64_t search_work(uint64_t searches, uint64_t n, uint64_t arcs) {
    unsigned __int128 w = (unsigned __int128)searches * (n + arcs);
    return w > UINT64_MAX ? UINT64_MAX : (uint64_t)w;
}

int64_t orc_threads_for(int64_t requested, int64_t items, uint64_t work) {
    return (int64_t)threads_for((uint64_t)requested, (uint64_t)items, work);
}

/* ---- sum_in_order: parallel.rs:379-503 --------------------------------- */

#define SUM_GROUPS 64

typedef struct {
    const OGraph *g;
    const Lengths *L;
    const int64_t *sources; /* NULL: item i is source i */
    uint64_t items, size, groups, window;
    /* shared */
    pthread_mutex_t mu;
    pthread_cond_t turn;
    uint64_t claim;
    uint64_t next;
    double **done;
    double **pool;
    uint64_t npool;
    double *bc;
    /* edge betweenness: accumulators of m doubles, arc -> edge map */
    int edge;
    const uint32_t *arc_edge;
} SumCtx;

static uint64_t acc_len(const SumCtx *c) { return c->edge ? c->g->m : c->g->n; }

static void merge_group(SumCtx *c, double *acc) {
    /* merge closures, betweenness.rs:91-96 and reach.rs:201-206 */
    uint64_t len = acc_len(c);
    for (uint64_t v = 0; v < len; v++) {
        c->bc[v] += acc[v];
        acc[v] = 0.0;
    }
}

static void work_group(SumCtx *c, Workspace *ws, double *acc, uint64_t grp) {
    uint64_t lo = grp * c->size, hi = (grp + 1) * c->size;
    if (hi > c->items) hi = c->items;
    ws->edge = c->edge;
    ws->arc_edge = c->arc_edge;
    if (c->edge) ws->eb = acc;
    else ws->bc = acc;
    for (uint64_t i = lo; i < hi; i++) {
        uint64_t s = c->sources ? (uint64_t)c->sources[i] : i;
        run_source(ws, c->g, c->L, s);
    }
    ws->bc = NULL;
    ws->eb = NULL;
}

static void *sum_worker(void *arg) {
    SumCtx *c = arg;
    Workspace ws;
    ws_init(&ws, c->g->n);
    for (;;) {
        pthread_mutex_lock(&c->mu);
        uint64_t grp = c->claim++;
        if (grp >= c->groups) { pthread_mutex_unlock(&c->mu); break; }
        while (grp >= c->next + c->window) pthread_cond_wait(&c->turn, &c->mu);
        double *acc = c->npool ? c->pool[--c->npool] : NULL;
        pthread_mutex_unlock(&c->mu);
        if (!acc) acc = calloc(acc_len(c) ? acc_len(c) : 1, sizeof(double));
        work_group(c, &ws, acc, grp);
        pthread_mutex_lock(&c->mu);
        c->done[grp] = acc;
        int merged = 0;
        while (c->next < c->groups && c->done[c->next]) {
            double *a = c->done[c->next];
            c->done[c->next] = NULL;
            merge_group(c, a);
            c->pool[c->npool++] = a;
            c->next++;
            merged = 1;
        }
        pthread_mutex_unlock(&c->mu);
        if (merged) pthread_cond_broadcast(&c->turn);
    }
    ws_free(&ws);
    return NULL;
}

static void sum_in_order(SumCtx *c, uint64_t threads) {
    if (c->items == 0) return;
    c->size = (c->items + SUM_GROUPS - 1) / SUM_GROUPS;
    c->groups = (c->items + c->size - 1) / c->size;
    threads = resolve_threads(threads, c->groups);
    if (threads == 1) {
        Workspace ws;
        ws_init(&ws, c->g->n);
        double *acc = calloc(acc_len(c) ? acc_len(c) : 1, sizeof(double));
        for (uint64_t grp = 0; grp < c->groups; grp++) {
            work_group(c, &ws, acc, grp);
            merge_group(c, acc);
        }
        free(acc);
        ws_free(&ws);
        return;
    }
    c->window = 2 * threads;
    pthread_mutex_init(&c->mu, NULL);
    pthread_cond_init(&c->turn, NULL);
    c->claim = 0;
    c->next = 0;
    c->done = calloc(c->groups, sizeof(double *));
    c->pool = calloc(c->groups + threads + 1, sizeof(double *));
    c->npool = 0;
    pthread_t *tid = malloc(threads * sizeof *tid);
    for (uint64_t t = 0; t < threads; t++) pthread_create(&tid[t], NULL, sum_worker, c);
    for (uint64_t t = 0; t < threads; t++) pthread_join(tid[t], NULL);
    for (uint64_t i = 0; i < c->npool; i++) free(c->pool[i]);
    free(c->pool);
    free(c->done);
    free(tid);
    pthread_mutex_destroy(&c->mu);
    pthread_cond_destroy(&c->turn);
}

/* ---- betweenness: betweenness.rs:63-115 -------------------------------- */

static void scale_scores(const OGraph *g, int normalized, double *bc) {
    uint64_t n = g->n;
    double scale = g->directed ? 1.0 : 0.5;
    if (normalized && n > 2) {
        double pairs = (double)((n - 1) * (n - 2));
        scale *= g->directed ? 1.0 / pairs : 2.0 / pairs;
    }
    if (scale != 1.0)
        for (uint64_t v = 0; v < n; v++) bc[v] *= scale;
}

static int check_kind(int32_t weights_as, char *err, size_t err_len) {
    if (weights_as == 0 || weights_as == 1) return ORC_OK;
    char msg[128];
    snprintf(msg, sizeof msg, "unknown weight kind code %d", weights_as);
    write_err(err, err_len, msg);
    return ORC_FAILED;
}

/* net_betweenness, lib.rs:214-244 -> betweenness(), betweenness.rs:63-115.
 * algorithm: 0 Auto, 1 Bfs, 2 Dial, 3 Dijkstra (Auto through the C ABI). */
int32_t orc_betweenness_alg(const OGraph *g, int32_t weighted, int32_t weights_as,
                            int32_t normalized, int32_t threads, int32_t algorithm, double *out,
                            char *err, size_t err_len) {
    if (!g) { write_err(err, err_len, "graph handle is null"); return ORC_FAILED; }
    if (check_kind(weights_as, err, err_len)) return ORC_FAILED;
    Lengths L;
    if (path_lengths(g, weighted, weights_as, algorithm, &L, err, err_len)) return ORC_FAILED;
    uint64_t n = g->n;
    memset(out, 0, n * sizeof *out);
    SumCtx c;
    memset(&c, 0, sizeof c);
    c.g = g;
    c.L = &L;
    c.items = n;
    c.bc = out;
    uint64_t t = threads_for(threads > 0 ? (uint64_t)threads : 0, n,
                             search_work(n, n, g->out.arcs));
    sum_in_order(&c, t);
    scale_scores(g, normalized, out);
    lengths_free(&L);
    return ORC_OK;
}

int32_t orc_betweenness(const OGraph *g, int32_t weighted, int32_t weights_as,
                        int32_t normalized, int32_t threads, double *out, char *err,
                        size_t err_len) {
    return orc_betweenness_alg(g, weighted, weights_as, normalized, threads, 0, out, err, err_len);
}

/* The same computation restricted to a list of k sources: the per-source
 * searches and the ordered group sum run over the list (items are list
 * positions), and the result gets the same scaling as betweenness(). It is
 * the comparator for sampled-source runs. */
int32_t orc_betweenness_sources(const OGraph *g, const int64_t *sources, int64_t k,
                                int32_t weighted, int32_t weights_as, int32_t normalized,
                                int32_t threads, double *out, char *err, size_t err_len) {
    if (!g) { write_err(err, err_len, "graph handle is null"); return ORC_FAILED; }
    if (check_kind(weights_as, err, err_len)) return ORC_FAILED;
    for (int64_t i = 0; i < k; i++) {
        if (sources[i] < 0 || (uint64_t)sources[i] >= g->n) {
            write_err(err, err_len, "source out of range");
            return ORC_FAILED;
        }
    }
    Lengths L;
    if (path_lengths(g, weighted, weights_as, 0, &L, err, err_len)) return ORC_FAILED;
    memset(out, 0, g->n * sizeof *out);
    SumCtx c;
    memset(&c, 0, sizeof c);
    c.g = g;
    c.L = &L;
    c.items = (uint64_t)k;
    c.sources = sources;
    c.bc = out;
    uint64_t t = threads_for(threads > 0 ? (uint64_t)threads : 0, (uint64_t)k,
                             search_work((uint64_t)k, g->n, g->out.arcs));
    sum_in_order(&c, t);
    scale_scores(g, normalized, out);
    lengths_free(&L);
    return ORC_OK;
}

/* arc_edges, reach.rs:234-253: the input edge behind every out-arc in CSR
 * order. Arcs are generated edge by edge (for undirected non-loop edges the
 * reverse right after) and stably sorted by (source, target), which is the
 * order Csr::build leaves them in. */
static int arc_cmp(const void *pa, const void *pb) {
    const uint64_t *a = pa, *b = pb;
    if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
    if (a[1] != b[1]) return a[1] < b[1] ? -1 : 1;
    return a[2] < b[2] ? -1 : (a[2] > b[2]);
}

static uint32_t *arc_edges(const OGraph *g) {
    uint64_t arcs = g->out.arcs, na = 0;
    uint64_t *t = malloc((arcs ? arcs : 1) * 4 * sizeof *t); /* src, tgt, gen, edge */
    for (uint64_t e = 0; e < g->m; e++) {
        uint64_t *r = t + 4 * na;
        r[0] = g->src[e]; r[1] = g->tgt[e]; r[2] = na; r[3] = e;
        na++;
        if (!g->directed && g->src[e] != g->tgt[e]) {
            r = t + 4 * na;
            r[0] = g->tgt[e]; r[1] = g->src[e]; r[2] = na; r[3] = e;
            na++;
        }
    }
    qsort(t, na, 4 * sizeof *t, arc_cmp);
    uint32_t *map = malloc((na ? na : 1) * sizeof *map);
    for (uint64_t i = 0; i < na; i++) map[i] = (uint32_t)t[4 * i + 3];
    free(t);
    return map;
}

/* net_edge_betweenness, crates/capi/src/roles.rs:77-100 -> edge_betweenness,
 * reach.rs:174-232: per-edge sums over all sources in 64 ordered groups,
 * then scale = (directed ? 1 : 0.5), divided by n(n - 1) as a double when
 * normalized (times 2 undirected; NaN below 2 nodes). out: m doubles in
 * input edge order. The _sources form runs over a list of k sources
 * (sources == NULL: all n nodes), items being list positions, with the
 * same scaling; it is the comparator for sampled-source runs. */
int32_t orc_edge_betweenness_sources(const OGraph *g, const int64_t *sources, int64_t k,
                                     int32_t weighted, int32_t weights_as, int32_t normalized,
                                     int32_t threads, double *out, char *err, size_t err_len) {
    if (!g) { write_err(err, err_len, "graph handle is null"); return ORC_FAILED; }
    if (check_kind(weights_as, err, err_len)) return ORC_FAILED;
    for (int64_t i = 0; sources && i < k; i++) {
        if (sources[i] < 0 || (uint64_t)sources[i] >= g->n) {
            write_err(err, err_len, "source out of range");
            return ORC_FAILED;
        }
    }
    Lengths L;
    if (path_lengths(g, weighted, weights_as, 0, &L, err, err_len)) return ORC_FAILED;
    uint64_t n = g->n, m = g->m;
    uint64_t items = sources ? (uint64_t)k : n;
    memset(out, 0, m * sizeof *out);
    uint32_t *map = arc_edges(g);
    SumCtx c;
    memset(&c, 0, sizeof c);
    c.g = g;
    c.L = &L;
    c.items = items;
    c.sources = sources;
    c.bc = out;
    c.edge = 1;
    c.arc_edge = map;
    uint64_t t = threads_for(threads > 0 ? (uint64_t)threads : 0, items,
                             search_work(items, n, g->out.arcs));
    sum_in_order(&c, t);
    double scale = g->directed ? 1.0 : 0.5;
    if (normalized) {
        double pairs = (double)n * ((double)n - 1.0);
        if (n < 2) scale = NAN;
        else if (g->directed) scale = scale / pairs;
        else scale = scale * 2.0 / pairs;
    }
    if (scale != 1.0)
        for (uint64_t e = 0; e < m; e++) out[e] *= scale;
    free(map);
    lengths_free(&L);
    return ORC_OK;
}

int32_t orc_edge_betweenness(const OGraph *g, int32_t weighted, int32_t weights_as,
                             int32_t normalized, int32_t threads, double *out, char *err,
                             size_t err_len) {
    return orc_edge_betweenness_sources(g, NULL, 0, weighted, weights_as, normalized, threads,
                                        out, err, err_len);
}

/* One search from s, exposing its state: hops (BFS) or integer distance
 * (Dial) in dist_u64 (UINT64_MAX when unreached), real distance in
 * dist_f64 (Dijkstra, INFINITY when unreached), path counts in sigma and
 * the dependency of every reached node in delta (0 elsewhere). Returns the
 * number of reached nodes in *reached and the length kind in *kind. */
int32_t orc_single_source(const OGraph *g, int32_t weighted, int32_t weights_as, int64_t s,
                          int64_t *reached, int32_t *kind, uint64_t *dist_u64, double *dist_f64,
                          double *sigma, double *delta, char *err, size_t err_len) {
    if (!g) { write_err(err, err_len, "graph handle is null"); return ORC_FAILED; }
    if (check_kind(weights_as, err, err_len)) return ORC_FAILED;
    if (s < 0 || (uint64_t)s >= g->n) { write_err(err, err_len, "source out of range"); return ORC_FAILED; }
    Lengths L;
    if (path_lengths(g, weighted, weights_as, 0, &L, err, err_len)) return ORC_FAILED;
    uint64_t n = g->n;
    Workspace ws;
    ws_init(&ws, n);
    for (uint64_t v = 0; v < n; v++) {
        delta[v] = 0.0;
        sigma[v] = 0.0;
        if (dist_u64) dist_u64[v] = UINT64_MAX;
        if (dist_f64) dist_f64[v] = INFINITY;
    }
    ws.delta_out = delta;
    switch (L.kind) {
    case LEN_UNIFORM: single_source_bfs(&ws, g, (uint64_t)s); break;
    case LEN_INTEGER: single_source_dial(&ws, g, L.ilens, L.max, (uint64_t)s); break;
    default: single_source_dijkstra(&ws, g, L.rlens, (uint64_t)s); break;
    }
    for (uint64_t i = 0; i < ws.norder; i++) {
        uint32_t v = ws.order[i];
        sigma[v] = ws.sigma[v];
        if (L.kind == LEN_UNIFORM && dist_u64) dist_u64[v] = ws.hops[v];
        if (L.kind == LEN_INTEGER && dist_u64) dist_u64[v] = ws.idist[v];
        if (L.kind == LEN_REAL && dist_f64) dist_f64[v] = ws.dist[v];
    }
    *reached = (int64_t)ws.norder;
    *kind = L.kind;
    ws_free(&ws);
    lengths_free(&L);
    return ORC_OK;
}

/* Test hook for sum_in_order's determinism: sums 1/(1+i) (x1e8 when
 * i % 3 == 0) over items in 64 ordered groups, as parallel.rs:567-585. */
typedef struct { uint64_t items, size, groups, window, claim, next, npool; double total;
                 double *done_v; int *done_f; pthread_mutex_t mu; pthread_cond_t turn; } TCtx;
static void *t_worker(void *arg) {
    TCtx *c = arg;
    for (;;) {
        pthread_mutex_lock(&c->mu);
        uint64_t g = c->claim++;
        if (g >= c->groups) { pthread_mutex_unlock(&c->mu); break; }
        while (g >= c->next + c->window) pthread_cond_wait(&c->turn, &c->mu);
        pthread_mutex_unlock(&c->mu);
        double acc = 0.0;
        uint64_t lo = g * c->size, hi = (g + 1) * c->size;
        if (hi > c->items) hi = c->items;
        for (uint64_t i = lo; i < hi; i++)
            acc += 1.0 / (1.0 + (double)i) * (i % 3 == 0 ? 1e8 : 1.0);
        pthread_mutex_lock(&c->mu);
        c->done_v[g] = acc;
        c->done_f[g] = 1;
        int merged = 0;
        while (c->next < c->groups && c->done_f[c->next]) {
            c->total += c->done_v[c->next];
            c->next++;
            merged = 1;
        }
        pthread_mutex_unlock(&c->mu);
        if (merged) pthread_cond_broadcast(&c->turn);
    }
    return NULL;
}

double orc_ordered_sum_test(uint64_t items, uint64_t threads) {
    if (items == 0) return 0.0;
    TCtx c;
    memset(&c, 0, sizeof c);
    c.items = items;
    c.size = (items + SUM_GROUPS - 1) / SUM_GROUPS;
    c.groups = (items + c.size - 1) / c.size;
    threads = resolve_threads(threads, c.groups);
    c.window = 2 * threads;
    c.done_v = calloc(c.groups, sizeof(double));
    c.done_f = calloc(c.groups, sizeof(int));
    pthread_mutex_init(&c.mu, NULL);
    pthread_cond_init(&c.turn, NULL);
    pthread_t *tid = malloc(threads * sizeof *tid);
    for (uint64_t t = 0; t < threads; t++) pthread_create(&tid[t], NULL, t_worker, &c);
    for (uint64_t t = 0; t < threads; t++) pthread_join(tid[t], NULL);
    free(tid); free(c.done_v); free(c.done_f);
    pthread_mutex_destroy(&c.mu);
    pthread_cond_destroy(&c.turn);
    return c.total;
}

uint64_t orc_available_cores(void) { return available_cores(); }

/* ---- betweenness for sources and targets: path_centrality.rs:421-689 --- */

/* PathGraph::new, paths/mod.rs:173-299 (paths follow arcs, mode Out): arcs
 * generated edge by edge (for undirected non-loop edges the reverse right
 * after), stably sorted by (tail, head); with collapse, one arc per
 * (tail, head), of the least length, the first edge of that length. */
typedef struct {
    uint64_t n, arcs;
    uint64_t *off;
    uint32_t *head, *tail, *edge;
    double *len;
    int unit, integral;
    double min_len, max_len;
} OPathGraph;

static void pg_free(OPathGraph *pg) {
    free(pg->off);
    free(pg->head);
    free(pg->tail);
    free(pg->edge);
    free(pg->len);
    memset(pg, 0, sizeof *pg);
}

static int pg_cmp(const void *pa, const void *pb) {
    const uint64_t *a = pa, *b = pb;
    for (int i = 0; i < 3; i++)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

/* path_graph, path_centrality.rs:198-216, with PathGraph::new's length
 * checks (paths/mod.rs:175-192) and require_positive (paths/mod.rs:416-425). */
static int pg_build(const OGraph *g, int weighted, int kind, int collapse, OPathGraph *pg,
                    char *err, size_t err_len) {
    char msg[512], num[64];
    memset(pg, 0, sizeof *pg);
    uint64_t n = g->n, m = g->m;
    double *elen = malloc((m ? m : 1) * sizeof *elen);
    for (uint64_t e = 0; e < m; e++) elen[e] = 1.0;
    if (weighted) {
        if (!g->w) {
            free(elen);
            write_err(err, err_len, "the graph has no edge weights");
            return ORC_FAILED;
        }
        for (uint64_t e = 0; e < m; e++) {
            if (kind == 1 && g->w[e] <= 0.0) {
                fmt_f64(g->w[e], num, sizeof num);
                snprintf(msg, sizeof msg,
                         "edge %llu has weight %s; weights read as strengths become lengths "
                         "1/w and must be strictly positive",
                         (unsigned long long)e, num);
                free(elen);
                write_err(err, err_len, msg);
                return ORC_FAILED;
            }
            elen[e] = kind == 1 ? 1.0 / g->w[e] : g->w[e];
        }
    }
    uint64_t *t = malloc((2 * m + 1) * 4 * sizeof *t); /* tail, head, gen, edge */
    uint64_t na = 0;
    for (uint64_t e = 0; e < m; e++) {
        uint64_t s = g->src[e], h = g->tgt[e];
        uint64_t *r = t + 4 * na;
        r[0] = s; r[1] = h; r[2] = na; r[3] = e;
        na++;
        if (!g->directed && s != h) {
            r = t + 4 * na;
            r[0] = h; r[1] = s; r[2] = na; r[3] = e;
            na++;
        }
    }
    qsort(t, na, 4 * sizeof *t, pg_cmp);
    pg->n = n;
    pg->head = malloc((na ? na : 1) * sizeof *pg->head);
    pg->tail = malloc((na ? na : 1) * sizeof *pg->tail);
    pg->edge = malloc((na ? na : 1) * sizeof *pg->edge);
    pg->len = malloc((na ? na : 1) * sizeof *pg->len);
    uint64_t k = 0;
    for (uint64_t i = 0; i < na; i++) {
        uint32_t tl = (uint32_t)t[4 * i], hd = (uint32_t)t[4 * i + 1];
        uint32_t e = (uint32_t)t[4 * i + 3];
        double len = elen[e];
        if (collapse && k > 0 && pg->tail[k - 1] == tl && pg->head[k - 1] == hd) {
            if (len < pg->len[k - 1]) {
                pg->len[k - 1] = len;
                pg->edge[k - 1] = e;
            }
            continue;
        }
        pg->head[k] = hd;
        pg->tail[k] = tl;
        pg->len[k] = len;
        pg->edge[k] = e;
        k++;
    }
    pg->arcs = k;
    free(t);
    free(elen);
    pg->off = calloc(n + 1, sizeof *pg->off);
    for (uint64_t a = 0; a < k; a++) pg->off[(uint64_t)pg->tail[a] + 1]++;
    for (uint64_t v = 0; v < n; v++) pg->off[v + 1] += pg->off[v];
    pg->unit = 1;
    pg->integral = 1;
    pg->min_len = INFINITY;
    pg->max_len = -INFINITY;
    for (uint64_t a = 0; a < k; a++) {
        double x = pg->len[a];
        if (x != 1.0) pg->unit = 0;
        if (!(x - trunc(x) == 0.0 && fabs(x) < 4503599627370496.0)) pg->integral = 0;
        if (x < pg->min_len) pg->min_len = x;
        if (x > pg->max_len) pg->max_len = x;
    }
    if (weighted && pg->min_len <= 0.0) {
        uint64_t a = 0;
        while (a < k && !(pg->len[a] <= 0.0)) a++;
        fmt_f64(pg->len[a], num, sizeof num);
        snprintf(msg, sizeof msg,
                 "edge %u has length %s; shortest-path betweenness needs strictly positive "
                 "lengths",
                 pg->edge[a], num);
        pg_free(pg);
        write_err(err, err_len, msg);
        return ORC_FAILED;
    }
    return ORC_OK;
}

/* Tie::equal, paths/mod.rs:96-104 (tie 0: relative 1e-10, 1: exact). */
static int tie_equal(int tie, double a, double b) {
    if (a == b) return 1;
    if (tie != 0 || !isfinite(a) || !isfinite(b)) return 0;
    double m = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
    return fabs(a - b) <= 1e-10 * m;
}

/* The shortest-path DAG from one source, path_centrality.rs:225-409. */
typedef struct {
    uint64_t n;
    double *dist, *sigma;
    uint32_t *rank, *order;
    uint64_t norder;
    U32Vec *preds;
    Heap heap;
    U32Vec *buckets;
    uint64_t nbuckets;
} ODag;

static void dag_init(ODag *d, uint64_t n) {
    memset(d, 0, sizeof *d);
    d->n = n;
    d->dist = malloc((n ? n : 1) * sizeof *d->dist);
    d->sigma = malloc((n ? n : 1) * sizeof *d->sigma);
    d->rank = malloc((n ? n : 1) * sizeof *d->rank);
    d->order = malloc((n ? n : 1) * sizeof *d->order);
    d->preds = calloc(n ? n : 1, sizeof *d->preds);
    d->heap.items = malloc((n ? n : 1) * sizeof *d->heap.items);
    d->heap.pos = malloc((n ? n : 1) * sizeof *d->heap.pos);
    for (uint64_t v = 0; v < n; v++) {
        d->dist[v] = INFINITY;
        d->sigma[v] = 0.0;
        d->rank[v] = U32_MAX_;
        d->heap.pos[v] = U32_MAX_;
    }
}

static void dag_free(ODag *d) {
    for (uint64_t v = 0; v < d->n; v++) free(d->preds[v].a);
    for (uint64_t b = 0; b < d->nbuckets; b++) free(d->buckets[b].a);
    free(d->buckets);
    free(d->dist);
    free(d->sigma);
    free(d->rank);
    free(d->order);
    free(d->preds);
    free(d->heap.items);
    free(d->heap.pos);
}

static void dag_reset(ODag *d) {
    for (uint64_t v = 0; v < d->n; v++) {
        d->dist[v] = INFINITY;
        d->sigma[v] = 0.0;
        d->rank[v] = U32_MAX_;
        d->preds[v].len = 0;
        d->heap.pos[v] = U32_MAX_;
    }
    d->norder = 0;
    d->heap.len = 0;
}

static double pred_sigma(const ODag *d, const OPathGraph *pg, uint32_t v) {
    double s = 0.0;
    for (uint64_t i = 0; i < d->preds[v].len; i++) s += d->sigma[pg->tail[d->preds[v].a[i]]];
    return s;
}

/* Dag::run without a cutoff: breadth-first (unit lengths), Dial's buckets
 * (integer lengths up to DIAL_MAX; buckets are stacks), or the 4-ary heap
 * with the tie rule. */
static void dag_run(ODag *d, const OPathGraph *pg, uint32_t s, int tie) {
    dag_reset(d);
    d->dist[s] = 0.0;
    d->sigma[s] = 1.0;
    if (pg->unit) {
        d->rank[s] = 0;
        d->order[d->norder++] = s;
        uint64_t head = 0;
        while (head < d->norder) {
            uint32_t v = d->order[head++];
            if (v != s) d->sigma[v] = pred_sigma(d, pg, v);
            double next = d->dist[v] + 1.0;
            for (uint64_t a = pg->off[v]; a < pg->off[v + 1]; a++) {
                uint32_t w = pg->head[a];
                if (d->rank[w] == U32_MAX_) {
                    d->dist[w] = next;
                    d->rank[w] = (uint32_t)d->norder;
                    d->order[d->norder++] = w;
                }
                if (d->dist[w] == next) vpush(&d->preds[w], (uint32_t)a);
            }
        }
        return;
    }
    if (!pg->unit && pg->integral && pg->min_len >= 0.0 && pg->max_len <= (double)DIAL_MAX) {
        uint64_t width = (uint64_t)(pg->max_len > 0.0 ? pg->max_len : 0.0) + 1;
        if (d->nbuckets != width) {
            for (uint64_t b = 0; b < d->nbuckets; b++) free(d->buckets[b].a);
            free(d->buckets);
            d->buckets = calloc(width, sizeof *d->buckets);
            d->nbuckets = width;
        }
        for (uint64_t b = 0; b < width; b++) d->buckets[b].len = 0;
        vpush(&d->buckets[0], s);
        uint64_t pending = 1, cur = 0;
        while (pending > 0) {
            U32Vec *bk = &d->buckets[cur % width];
            if (bk->len == 0) {
                cur++;
                continue;
            }
            uint32_t v = bk->a[--bk->len];
            pending--;
            if (d->rank[v] != U32_MAX_ || d->dist[v] != (double)cur) continue;
            d->rank[v] = (uint32_t)d->norder;
            d->order[d->norder++] = v;
            if (v != s) d->sigma[v] = pred_sigma(d, pg, v);
            for (uint64_t a = pg->off[v]; a < pg->off[v + 1]; a++) {
                uint32_t w = pg->head[a];
                if (d->rank[w] != U32_MAX_) continue;
                uint64_t alt = cur + (uint64_t)pg->len[a];
                double altf = (double)alt;
                double dw = d->dist[w];
                if (isinf(dw) || altf < dw) {
                    d->dist[w] = altf;
                    d->preds[w].len = 0;
                    vpush(&d->preds[w], (uint32_t)a);
                    vpush(&d->buckets[alt % width], w);
                    pending++;
                } else if (altf == dw) {
                    vpush(&d->preds[w], (uint32_t)a);
                }
            }
        }
        return;
    }
    h_push_or_decrease(&d->heap, s, 0.0);
    HItem top;
    while (h_pop(&d->heap, &top)) {
        uint32_t v = top.v;
        double dv = top.d;
        d->rank[v] = (uint32_t)d->norder;
        d->order[d->norder++] = v;
        if (v != s) d->sigma[v] = pred_sigma(d, pg, v);
        for (uint64_t a = pg->off[v]; a < pg->off[v + 1]; a++) {
            uint32_t w = pg->head[a];
            if (d->rank[w] != U32_MAX_) continue;
            double alt = dv + pg->len[a];
            double cur = d->dist[w];
            if (isinf(cur)) {
                d->dist[w] = alt;
                vpush(&d->preds[w], (uint32_t)a);
                h_push_or_decrease(&d->heap, w, alt);
            } else if (tie_equal(tie, alt, cur)) {
                vpush(&d->preds[w], (uint32_t)a);
            } else if (alt < cur) {
                d->dist[w] = alt;
                d->preds[w].len = 0;
                vpush(&d->preds[w], (uint32_t)a);
                h_push_or_decrease(&d->heap, w, alt);
            }
        }
    }
}

/* networkx_rescale, path_centrality.rs:596-649. sampled < 0: no sample. */
static void nx_rescale(double *values, uint64_t len, uint64_t n, int normalized, int directed,
                       int endpoints, int64_t sampled, const uint8_t *is_source) {
    double big_n = endpoints ? (double)n : (double)n - 1.0;
    if (big_n < 2.0) return;
    double correction = directed ? 1.0 : 2.0;
    double k_source = sampled >= 0 ? (double)sampled : big_n;
    if (sampled < 0 || endpoints) {
        double scale = normalized ? 1.0 / (k_source * (big_n - 1.0))
                                  : big_n / (k_source * correction);
        if (scale != 1.0)
            for (uint64_t v = 0; v < len; v++) values[v] *= scale;
        return;
    }
    double ss, so;
    if (normalized) {
        ss = k_source > 1.0 ? 1.0 / ((k_source - 1.0) * (big_n - 1.0)) : NAN;
        so = 1.0 / (k_source * (big_n - 1.0));
    } else {
        ss = k_source > 1.0 ? big_n / ((k_source - 1.0) * correction) : NAN;
        so = big_n / (k_source * correction);
    }
    for (uint64_t v = 0; v < len; v++) values[v] *= (is_source && is_source[v]) ? ss : so;
}

/* share_among_parallel, path_centrality.rs:651-689: each adjacent pair's
 * value is shared equally by its parallel edges of least length; other
 * parallel edges and loops get 0. Groups by (a, b) (sorted undirected). */
static void share_among_parallel(const OGraph *g, const double *lengths, double *values) {
    uint64_t m = g->m;
    uint64_t *t = malloc((m ? m : 1) * 3 * sizeof *t);
    for (uint64_t e = 0; e < m; e++) {
        uint64_t a = g->src[e], b = g->tgt[e];
        if (!g->directed && a > b) { uint64_t x = a; a = b; b = x; }
        t[3 * e] = a; t[3 * e + 1] = b; t[3 * e + 2] = e;
    }
    qsort(t, m, 3 * sizeof *t, pg_cmp);
    double *out = calloc(m ? m : 1, sizeof *out);
    for (uint64_t i = 0; i < m;) {
        uint64_t j = i;
        while (j < m && t[3 * j] == t[3 * i] && t[3 * j + 1] == t[3 * i + 1]) j++;
        if (t[3 * i] != t[3 * i + 1]) {
            double total = 0.0, shortest = INFINITY;
            for (uint64_t q = i; q < j; q++) total += values[t[3 * q + 2]];  /* input order */
            for (uint64_t q = i; q < j; q++)
                if (lengths[t[3 * q + 2]] < shortest) shortest = lengths[t[3 * q + 2]];
            uint64_t ties = 0;
            for (uint64_t q = i; q < j; q++) ties += lengths[t[3 * q + 2]] == shortest;
            for (uint64_t q = i; q < j; q++)
                if (lengths[t[3 * q + 2]] == shortest) out[t[3 * q + 2]] = total / (double)ties;
        }
        i = j;
    }
    memcpy(values, out, m * sizeof *values);
    free(out);
    free(t);
}

/* net_betweenness_subset, crates/capi/src/centrality.rs:89-144 ->
 * betweenness_subset, path_centrality.rs:421-593. sources / targets: NULL
 * (or a negative count) for every node. rescale 0 unordered, 1 NetworkX,
 * 2 NetworkX subset; parallel 0 distinct, 1 collapse; edge_split 0
 * proportional, 1 NetworkX; tie 0 relative, 1 exact. Sources are summed
 * in 64 ordered groups (sum_in_order), single-threaded here: the result
 * does not depend on the thread count. nodes: n doubles; edges: NULL or m. */
int32_t orc_betweenness_subset(const OGraph *g, const int64_t *sources, int64_t n_sources,
                               const int64_t *targets, int64_t n_targets, int32_t weighted,
                               int32_t weights_as, int32_t endpoints, int32_t normalized,
                               int32_t rescale, int32_t parallel, int32_t edge_split,
                               int32_t tie, double *nodes, double *edges, char *err,
                               size_t err_len) {
    char msg[256];
    if (!g) { write_err(err, err_len, "graph handle is null"); return ORC_FAILED; }
    uint64_t n = g->n, m = g->m;
    const int64_t *lists[2] = {sources, targets};
    int64_t counts[2] = {n_sources, n_targets};
    for (int l = 0; l < 2; l++) {
        if (!lists[l] || counts[l] < 0) continue;
        for (int64_t i = 0; i < counts[l]; i++) {
            int64_t v = lists[l][i];
            if (v < 0 || (uint64_t)v >= n) {
                snprintf(msg, sizeof msg,
                         "node index %lld is out of range for a network with %llu nodes",
                         (long long)v, (unsigned long long)n);
                write_err(err, err_len, msg);
                return ORC_FAILED;
            }
        }
    }
    if (check_kind(weights_as, err, err_len)) return ORC_FAILED;
    if (rescale < 0 || rescale > 2) {
        snprintf(msg, sizeof msg, "unknown rescaling code %d", rescale);
        write_err(err, err_len, msg);
        return ORC_FAILED;
    }
    if (parallel < 0 || parallel > 1) {
        snprintf(msg, sizeof msg, "unknown code %d for parallel edges", parallel);
        write_err(err, err_len, msg);
        return ORC_FAILED;
    }
    if (edge_split < 0 || edge_split > 1) {
        snprintf(msg, sizeof msg, "unknown edge split code %d", edge_split);
        write_err(err, err_len, msg);
        return ORC_FAILED;
    }
    if (tie < 0 || tie > 1) {
        snprintf(msg, sizeof msg, "unknown tie code %d", tie);
        write_err(err, err_len, msg);
        return ORC_FAILED;
    }
    int collapse = parallel == 1;
    OPathGraph pg;
    if (pg_build(g, weighted, weights_as, collapse, &pg, err, err_len)) return ORC_FAILED;
    int all_src = !sources || n_sources < 0;
    uint64_t ns = all_src ? n : (uint64_t)n_sources;
    uint32_t *src = malloc((ns ? ns : 1) * sizeof *src);
    for (uint64_t i = 0; i < ns; i++) src[i] = all_src ? (uint32_t)i : (uint32_t)sources[i];
    uint8_t *is_target = calloc(n ? n : 1, 1);
    if (!targets || n_targets < 0) {
        for (uint64_t v = 0; v < n; v++) is_target[v] = 1;
    } else {
        for (int64_t i = 0; i < n_targets; i++) is_target[targets[i]] = 1;
    }
    int quirk = edge_split == 1;
    double *acc_n = calloc(n ? n : 1, sizeof *acc_n), *acc_e = calloc(m ? m : 1, sizeof *acc_e);
    double *tot_n = calloc(n ? n : 1, sizeof *tot_n), *tot_e = calloc(m ? m : 1, sizeof *tot_e);
    double *delta = calloc(n ? n : 1, sizeof *delta), *delta_e = calloc(n ? n : 1, sizeof *delta_e);
    ODag d;
    dag_init(&d, n);
    uint64_t size = (ns + SUM_GROUPS - 1) / SUM_GROUPS;
    uint64_t groups = size ? (ns + size - 1) / size : 0;
    for (uint64_t grp = 0; grp < groups; grp++) {
        uint64_t lo = grp * size, hi = lo + size < ns ? lo + size : ns;
        for (uint64_t k = lo; k < hi; k++) {
            uint32_t s = src[k];
            dag_run(&d, &pg, s, tie);
            uint64_t reached_targets = 0;
            for (uint64_t i = d.norder; i-- > 0;) {
                uint32_t w = d.order[i];
                double tw = (is_target[w] && w != s) ? 1.0 : 0.0;
                double coeff = (tw + delta[w]) / d.sigma[w];
                double np = (double)d.preds[w].len;
                for (uint64_t q = 0; q < d.preds[w].len; q++) {
                    uint32_t a = d.preds[w].a[q];
                    uint32_t v = pg.tail[a];
                    double c = d.sigma[v] * coeff;
                    delta[v] += c;
                    double ce = c;
                    if (quirk) {
                        ce = is_target[w] ? d.sigma[v] / d.sigma[w] * (1.0 + delta_e[w])
                                          : delta_e[w] / np;
                        delta_e[v] += ce;
                    }
                    acc_e[pg.edge[a]] += ce;
                }
                if (w != s) {
                    acc_n[w] += delta[w];
                    if (endpoints && tw > 0.0) {
                        acc_n[w] += 1.0;
                        reached_targets++;
                    }
                }
            }
            if (endpoints) acc_n[s] += (double)reached_targets;
            for (uint64_t i = 0; i < d.norder; i++) {
                delta[d.order[i]] = 0.0;
                delta_e[d.order[i]] = 0.0;
            }
        }
        for (uint64_t v = 0; v < n; v++) { tot_n[v] += acc_n[v]; acc_n[v] = 0.0; }
        for (uint64_t e = 0; e < m; e++) { tot_e[e] += acc_e[e]; acc_e[e] = 0.0; }
    }
    /* scaling, path_centrality.rs:527-574 */
    uint8_t *source_set = calloc(n ? n : 1, 1);
    for (uint64_t i = 0; i < ns; i++) source_set[src[i]] = 1;
    int64_t sampled = all_src ? -1 : (int64_t)ns;
    if (rescale == 0) {
        double half = g->directed ? 1.0 : 0.5;
        double np_ = 1.0, ep_ = 1.0;
        if (normalized) {
            double nf = (double)n;
            np_ = endpoints ? nf * (nf - 1.0) : (nf - 1.0) * (nf - 2.0);
            ep_ = nf * (nf - 1.0);
        }
        double fn = np_ > 0.0 ? half / np_ : half;
        double fe = ep_ > 0.0 ? half / ep_ : half;
        for (uint64_t v = 0; v < n; v++) tot_n[v] *= fn;
        for (uint64_t e = 0; e < m; e++) tot_e[e] *= fe;
    } else if (rescale == 1) {
        nx_rescale(tot_n, n, n, normalized, g->directed, endpoints, sampled, source_set);
        nx_rescale(tot_e, m, n, normalized, g->directed, 1, sampled, NULL);
    } else {
        nx_rescale(tot_n, n, n, normalized, g->directed, 0, -1, NULL);
        nx_rescale(tot_e, m, n, normalized, g->directed, 1, -1, NULL);
    }
    if (collapse) {
        double *lengths = malloc((m ? m : 1) * sizeof *lengths);
        for (uint64_t e = 0; e < m; e++)
            lengths[e] = (weighted && g->w) ? (weights_as == 1 ? 1.0 / g->w[e] : g->w[e]) : 1.0;
        share_among_parallel(g, lengths, tot_e);
        free(lengths);
    }
    memcpy(nodes, tot_n, n * sizeof *nodes);
    if (edges) memcpy(edges, tot_e, m * sizeof *edges);
    dag_free(&d);
    pg_free(&pg);
    free(src); free(is_target); free(source_set);
    free(acc_n); free(acc_e); free(tot_n); free(tot_e); free(delta); free(delta_e);
    return ORC_OK;
}

/* ---- closeness and harmonic centrality: closeness.rs:131-481 ----------- */

/* Profile, closeness.rs:212-223: other nodes reached, sum of lengths, sum
 * of reciprocal lengths, longest length. */
typedef struct { uint64_t reached; double total, harmonic, longest; } OProfile;

/* MultiBfs's per-level bookkeeping (closeness.rs:268-293), one source:
 * reached += c, total += c * level (an integer), harmonic += c / level. */
static OProfile profile_bfs(const OGraph *g, uint32_t s, uint32_t *hops, uint32_t *queue) {
    const OCsr *adj = &g->out;
    OProfile p = {0, 0.0, 0.0, 0.0};
    uint64_t total = 0, head = 0, tail = 0;
    hops[s] = 0;
    queue[tail++] = s;
    uint64_t level_start = tail;
    uint32_t level = 0;
    uint64_t count = 0;
    while (head < tail) {
        uint32_t v = queue[head++];
        for (uint64_t a = adj->offsets[v]; a < adj->offsets[v + 1]; a++) {
            uint32_t w = adj->targets[a];
            if (hops[w] != U32_MAX_) continue;
            hops[w] = hops[v] + 1;
            queue[tail++] = w;
        }
        if (head == level_start) {
            /* level `level` + 1 is complete */
            count = tail - level_start;
            level++;
            if (count) {
                p.reached += count;
                total += count * (uint64_t)level;
                p.harmonic += (double)count / (double)level;
                p.longest = (double)level;
            }
            level_start = tail;
        }
    }
    p.total = (double)total;
    for (uint64_t i = 0; i < tail; i++) hops[queue[i]] = U32_MAX_;
    return p;
}

/* Search::dial, closeness.rs:364-425: buckets are stacks. */
static OProfile profile_dial(const OGraph *g, const uint32_t *lens, uint32_t max, uint32_t s,
                             uint64_t *idist, uint8_t *done, U32Vec *buckets, U32Vec *touched) {
    const OCsr *adj = &g->out;
    uint64_t width = (uint64_t)max + 1;
    touched->len = 0;
    vpush(touched, s);
    idist[s] = 0;
    for (uint64_t b = 0; b < width; b++) buckets[b].len = 0;
    vpush(&buckets[0], s);
    uint64_t pending = 1, cur = 0, reached = 0, total = 0, longest = 0;
    double harmonic = 0.0;
    while (pending > 0) {
        U32Vec *bk = &buckets[cur % width];
        if (bk->len == 0) { cur++; continue; }
        uint32_t v = bk->a[--bk->len];
        pending--;
        if (done[v] || idist[v] != cur) continue;
        done[v] = 1;
        if (v != s) {
            reached++;
            total += cur;
            harmonic += 1.0 / (double)cur;
            if (cur > longest) longest = cur;
        }
        for (uint64_t a = adj->offsets[v]; a < adj->offsets[v + 1]; a++) {
            uint32_t w = adj->targets[a];
            if (done[w]) continue;
            uint64_t alt = cur + lens[a];
            if (idist[w] == UINT64_MAX) vpush(touched, w);
            if (alt < idist[w]) {
                idist[w] = alt;
                vpush(&buckets[alt % width], w);
                pending++;
            }
        }
    }
    for (uint64_t i = 0; i < touched->len; i++) {
        idist[touched->a[i]] = UINT64_MAX;
        done[touched->a[i]] = 0;
    }
    OProfile p = {reached, (double)total, harmonic, (double)longest};
    return p;
}

/* Search::dijkstra, closeness.rs:427-470 (the 4-ary heap, no tie rule). */
static OProfile profile_dijkstra(const OGraph *g, const double *lens, uint32_t s, double *dist,
                                 uint8_t *done, Heap *heap, U32Vec *touched) {
    const OCsr *adj = &g->out;
    touched->len = 0;
    vpush(touched, s);
    dist[s] = 0.0;
    h_push_or_decrease(heap, s, 0.0);
    OProfile p = {0, 0.0, 0.0, 0.0};
    HItem top;
    while (h_pop(heap, &top)) {
        uint32_t v = top.v;
        double d = top.d;
        done[v] = 1;
        if (v != s) {
            p.reached++;
            p.total += d;
            p.harmonic += 1.0 / d;
            if (d > p.longest) p.longest = d;
        }
        for (uint64_t a = adj->offsets[v]; a < adj->offsets[v + 1]; a++) {
            uint32_t w = adj->targets[a];
            if (done[w]) continue;
            double alt = d + lens[a];
            if (isinf(dist[w])) vpush(touched, w);
            if (alt < dist[w]) {
                dist[w] = alt;
                h_push_or_decrease(heap, w, alt);
            }
        }
    }
    for (uint64_t i = 0; i < touched->len; i++) {
        dist[touched->a[i]] = INFINITY;
        done[touched->a[i]] = 0;
    }
    return p;
}

/* net_closeness, crates/capi/src/topography.rs:167-199 -> closeness /
 * harmonic_centrality (closeness.rs:131-209) over distance_profiles
 * (closeness.rs:240-345). mode: 1 out, 2 in, 3 all; unreachable: 0 ignore,
 * 1 Wasserman-Faust, 2 zero, 3 max plus one. */
int32_t orc_closeness(const OGraph *g0, int32_t harmonic, int32_t mode, int32_t weighted,
                      int32_t weights_as, int32_t normalized, int32_t unreachable,
                      double *out, char *err, size_t err_len) {
    char msg[256];
    if (!g0) { write_err(err, err_len, "graph handle is null"); return ORC_FAILED; }
    if (mode < 1 || mode > 3) {
        snprintf(msg, sizeof msg, "unknown degree mode code %d", mode);
        write_err(err, err_len, msg);
        return ORC_FAILED;
    }
    if (check_kind(weights_as, err, err_len)) return ORC_FAILED;
    if (unreachable < 0 || unreachable > 3) {
        snprintf(msg, sizeof msg, "unknown code %d for unreachable nodes", unreachable);
        write_err(err, err_len, msg);
        return ORC_FAILED;
    }
    if (!harmonic && unreachable == 1 && !normalized) {
        write_err(err, err_len,
                  "the Wasserman-Faust closeness index is defined only in normalized form");
        return ORC_FAILED;
    }
    if (!harmonic && unreachable == 3 && weighted) {
        write_err(err, err_len,
                  "closeness with unreachable lengths set to the longest length plus 1 is "
                  "defined only for unweighted networks");
        return ORC_FAILED;
    }
    /* oriented, closeness.rs:225-238: paths along in-arcs, or ignoring
     * direction, as the out-paths of a derived graph */
    const OGraph *g = g0;
    OGraph *h = NULL;
    if (g0->directed && mode != 1) {
        uint64_t m = g0->m;
        int64_t *s = malloc((m ? m : 1) * sizeof *s), *t = malloc((m ? m : 1) * sizeof *t);
        for (uint64_t e = 0; e < m; e++) {
            s[e] = mode == 2 ? g0->tgt[e] : g0->src[e];
            t[e] = mode == 2 ? g0->src[e] : g0->tgt[e];
        }
        int32_t rc = orc_graph_new((int64_t)g0->n, s, t, g0->w, (int64_t)m, mode == 2, &h, err,
                                   err_len);
        free(s);
        free(t);
        if (rc) return ORC_FAILED;
        g = h;
    }
    Lengths L;
    if (path_lengths(g, weighted, weights_as, 0, &L, err, err_len)) {
        orc_graph_free(h);
        return ORC_FAILED;
    }
    uint64_t n = g->n;
    OProfile *prof = calloc(n ? n : 1, sizeof *prof);
    uint32_t *hops = malloc((n ? n : 1) * sizeof *hops), *queue = malloc((n ? n : 1) * sizeof *queue);
    uint64_t *idist = malloc((n ? n : 1) * sizeof *idist);
    double *dist = malloc((n ? n : 1) * sizeof *dist);
    uint8_t *done = calloc(n ? n : 1, 1);
    for (uint64_t v = 0; v < n; v++) {
        hops[v] = U32_MAX_;
        idist[v] = UINT64_MAX;
        dist[v] = INFINITY;
    }
    U32Vec touched = {0};
    U32Vec *buckets = L.kind == LEN_INTEGER ? calloc((uint64_t)L.max + 1, sizeof *buckets) : NULL;
    Heap heap;
    heap.items = malloc((n ? n : 1) * sizeof *heap.items);
    heap.pos = malloc((n ? n : 1) * sizeof *heap.pos);
    heap.len = 0;
    for (uint64_t v = 0; v < n; v++) heap.pos[v] = U32_MAX_;
    for (uint64_t s = 0; s < n; s++) {
        if (L.kind == LEN_UNIFORM) prof[s] = profile_bfs(g, (uint32_t)s, hops, queue);
        else if (L.kind == LEN_INTEGER)
            prof[s] = profile_dial(g, L.ilens, L.max, (uint32_t)s, idist, done, buckets, &touched);
        else prof[s] = profile_dijkstra(g, L.rlens, (uint32_t)s, dist, done, &heap, &touched);
    }
    /* uniform weights: hop counts times the common length (closeness.rs:325-344) */
    if (L.kind == LEN_UNIFORM && weighted && g->out.arcs > 0 && g->out.weights) {
        double w = g->out.weights[0];
        double unit = weights_as == 1 ? 1.0 / w : w;
        if (unit != 1.0)
            for (uint64_t s = 0; s < n; s++) {
                prof[s].total *= unit;
                prof[s].harmonic /= unit;
                prof[s].longest *= unit;
            }
    }
    if (harmonic) {
        double scale = normalized && n > 1 ? 1.0 / (double)(n - 1) : 1.0;
        for (uint64_t v = 0; v < n; v++) out[v] = prof[v].harmonic * scale;
    } else {
        double longest = 0.0;
        for (uint64_t v = 0; v < n; v++) if (prof[v].longest > longest) longest = prof[v].longest;
        double others = n > 0 ? (double)(n - 1) : 0.0;
        for (uint64_t v = 0; v < n; v++) {
            OProfile *p = &prof[v];
            double r = (double)p->reached, x;
            switch (unreachable) {
            case 0:
                x = p->reached == 0 ? NAN : (normalized ? r / p->total : 1.0 / p->total);
                break;
            case 1:
                x = p->reached == 0 ? 0.0 : (r / p->total) * (r / others);
                break;
            case 2:
                x = n < 2 ? NAN
                    : p->reached < n - 1 ? 0.0
                    : normalized ? others / p->total : 1.0 / p->total;
                break;
            default: {
                if (n < 2) { x = NAN; break; }
                double total = p->total + (others - r) * (longest + 1.0);
                x = normalized ? others / total : 1.0 / total;
            }
            }
            out[v] = x;
        }
    }
    if (buckets) {
        for (uint64_t b = 0; b <= L.max; b++) free(buckets[b].a);
        free(buckets);
    }
    free(touched.a);
    free(heap.items);
    free(heap.pos);
    free(prof); free(hops); free(queue); free(idist); free(dist); free(done);
    lengths_free(&L);
    orc_graph_free(h);
    return ORC_OK;
}

/* ---- sampled distance profiles (timing baseline for closeness) --------- */

/* The profiles of closeness.rs:240-345 for a list of sources, along the
 * out-arcs of g, `threads` workers (0: every core) each searching every
 * threads-th source with private scratch. */
typedef struct {
    const OGraph *g;
    const Lengths *L;
    const int64_t *src;
    uint64_t k, stride, first;
    uint64_t *reached;
    double *total, *harmonic, *longest;
} ProfJob;

static void *profile_worker(void *arg) {
    ProfJob *j = arg;
    const OGraph *g = j->g;
    uint64_t n = g->n;
    uint32_t *hops = malloc((n ? n : 1) * sizeof *hops), *queue = malloc((n ? n : 1) * sizeof *queue);
    uint64_t *idist = malloc((n ? n : 1) * sizeof *idist);
    double *dist = malloc((n ? n : 1) * sizeof *dist);
    uint8_t *done = calloc(n ? n : 1, 1);
    for (uint64_t v = 0; v < n; v++) {
        hops[v] = U32_MAX_;
        idist[v] = UINT64_MAX;
        dist[v] = INFINITY;
    }
    U32Vec touched = {0};
    U32Vec *buckets = j->L->kind == LEN_INTEGER ? calloc((uint64_t)j->L->max + 1, sizeof *buckets)
                                                 : NULL;
    Heap heap;
    heap.items = malloc((n ? n : 1) * sizeof *heap.items);
    heap.pos = malloc((n ? n : 1) * sizeof *heap.pos);
    heap.len = 0;
    for (uint64_t v = 0; v < n; v++) heap.pos[v] = U32_MAX_;
    for (uint64_t i = j->first; i < j->k; i += j->stride) {
        uint32_t s = (uint32_t)j->src[i];
        OProfile p;
        if (j->L->kind == LEN_UNIFORM) p = profile_bfs(g, s, hops, queue);
        else if (j->L->kind == LEN_INTEGER)
            p = profile_dial(g, j->L->ilens, j->L->max, s, idist, done, buckets, &touched);
        else p = profile_dijkstra(g, j->L->rlens, s, dist, done, &heap, &touched);
        j->reached[i] = p.reached;
        j->total[i] = p.total;
        j->harmonic[i] = p.harmonic;
        j->longest[i] = p.longest;
    }
    if (buckets) {
        for (uint64_t b = 0; b <= j->L->max; b++) free(buckets[b].a);
        free(buckets);
    }
    free(touched.a);
    free(heap.items);
    free(heap.pos);
    free(hops); free(queue); free(idist); free(dist); free(done);
    return NULL;
}

int32_t orc_distance_profiles(const OGraph *g, int32_t weighted, int32_t weights_as,
                              const int64_t *sources, int64_t k, int32_t threads,
                              uint64_t *reached, double *total, double *harmonic,
                              double *longest, char *err, size_t err_len) {
    if (!g) { write_err(err, err_len, "graph handle is null"); return ORC_FAILED; }
    if (check_kind(weights_as, err, err_len)) return ORC_FAILED;
    for (int64_t i = 0; i < k; i++)
        if (sources[i] < 0 || (uint64_t)sources[i] >= g->n) {
            write_err(err, err_len, "source out of range");
            return ORC_FAILED;
        }
    Lengths L;
    if (path_lengths(g, weighted, weights_as, 0, &L, err, err_len)) return ORC_FAILED;
    uint64_t t = threads > 0 ? (uint64_t)threads : available_cores();
    if (t > (uint64_t)(k > 0 ? k : 1)) t = (uint64_t)(k > 0 ? k : 1);
    ProfJob *jobs = calloc(t, sizeof *jobs);
    pthread_t *tid = malloc(t * sizeof *tid);
    for (uint64_t i = 0; i < t; i++) {
        jobs[i] = (ProfJob){g, &L, sources, (uint64_t)k, t, i, reached, total, harmonic, longest};
        pthread_create(&tid[i], NULL, profile_worker, &jobs[i]);
    }
    for (uint64_t i = 0; i < t; i++) pthread_join(tid[i], NULL);
    free(jobs);
    free(tid);
    lengths_free(&L);
    return ORC_OK;
}

/* ---- centralization and two-mode centrality ----------------------------- */

/* simplify, topography.rs:105-140: loops dropped, one arc per ordered pair
 * (unordered when the result is undirected, smaller id first), the largest
 * weight of a pair kept, arcs sorted by (source, target). */
typedef struct { uint32_t s, t; double x; } SArc;

static int sarc_cmp(const void *a, const void *b) {
    const SArc *p = a, *q = b;
    if (p->s != q->s) return p->s < q->s ? -1 : 1;
    if (p->t != q->t) return p->t < q->t ? -1 : 1;
    return 0;
}

static int32_t simplify_max(const OGraph *g, int undirected, OGraph **out, char *err,
                            size_t err_len) {
    int directed = g->directed && !undirected;
    uint64_t m = g->m, k = 0;
    SArc *a = malloc((m ? m : 1) * sizeof *a);
    for (uint64_t e = 0; e < m; e++) {
        uint32_t s = g->src[e], t = g->tgt[e];
        if (s == t) continue;
        double x = g->w ? g->w[e] : 1.0;
        if (directed || s < t) a[k++] = (SArc){s, t, x};
        else a[k++] = (SArc){t, s, x};
    }
    qsort(a, k, sizeof *a, sarc_cmp);
    uint64_t mm = 0;
    for (uint64_t i = 0; i < k; i++) {
        if (mm > 0 && a[mm - 1].s == a[i].s && a[mm - 1].t == a[i].t) {
            if (a[i].x > a[mm - 1].x || isnan(a[mm - 1].x)) a[mm - 1].x = a[i].x;  /* f64::max */
        } else {
            a[mm++] = a[i];
        }
    }
    int64_t *s = malloc((mm ? mm : 1) * sizeof *s), *t = malloc((mm ? mm : 1) * sizeof *t);
    double *w = g->w ? malloc((mm ? mm : 1) * sizeof *w) : NULL;
    for (uint64_t i = 0; i < mm; i++) {
        s[i] = a[i].s;
        t[i] = a[i].t;
        if (w) w[i] = a[i].x;
    }
    int32_t rc = orc_graph_new((int64_t)g->n, s, t, w, (int64_t)mm, directed, out, err, err_len);
    free(a); free(s); free(t); free(w);
    return rc;
}

/* degree, degree.rs:41-74 (mode 1 out, 2 in, 3 all). */
static void degree_scores(const OGraph *g, int32_t mode, int loops, double *out) {
    for (uint64_t v = 0; v < g->n; v++) out[v] = 0.0;
    for (uint64_t e = 0; e < g->m; e++) {
        uint32_t s = g->src[e], t = g->tgt[e];
        if (s == t && !loops) continue;
        if (!g->directed) {
            out[s] += 1.0;
            out[t] += 1.0;
            continue;
        }
        if (mode != 2) out[s] += 1.0;
        if (mode != 1) out[t] += 1.0;
    }
}

/* centralization_max, topography.rs:504-568 (measure 0 degree, 1
 * betweenness, 2 closeness; convention 0 igraph, 2 sna, 3 UCINET). */
static double centralization_max(uint64_t n0, int directed, int32_t measure, int32_t mode,
                                 int loops, int32_t convention) {
    if (n0 == 0) return NAN;
    double n = (double)n0;
    if (!directed) mode = 3;
    loops = loops && convention == 0;
    if (measure == 0) {
        if (!directed) return loops ? (n - 1.0) * n : (n - 1.0) * (n - 2.0);
        if (mode == 3) return loops ? 2.0 * (n - 1.0) * (n - 1.0) : 2.0 * (n - 1.0) * (n - 2.0);
        return loops ? (n - 1.0) * n : (n - 1.0) * (n - 1.0);
    }
    if (measure == 1) {
        double t = (n - 1.0) * (n - 1.0) * (n - 2.0);
        return directed ? t : t / 2.0;
    }
    if (convention == 3 || mode == 3) return (n - 1.0) * (n - 2.0) / (2.0 * n - 3.0);
    return (n - 1.0) * (1.0 - 1.0 / n);
}

/* freeman_index, topography.rs:451-475: n * max - sum (in node order),
 * NaN when a score is NaN, over the maximum when normalized. */
static double freeman_index(uint64_t n, int directed, const double *scores, int32_t measure,
                            int32_t mode, int loops, int normalized, int32_t convention) {
    if (n == 0) return NAN;
    double max = -INFINITY, total = 0.0;
    int any_nan = 0;
    for (uint64_t v = 0; v < n; v++) {
        double x = scores[v];
        if (isnan(x)) any_nan = 1;
        else if (x > max) max = x;
        total += x;
    }
    if (any_nan) max = NAN;
    double raw = (double)n * max - total;
    if (!normalized) return raw;
    return raw / centralization_max(n, directed, measure, mode, loops, convention);
}

/* net_centralization, crates/capi/src/topography.rs:134-159 ->
 * centralization (topography.rs:376-449). measure 0 degree, 1
 * betweenness, 2 closeness (3, eigenvector, is checked and then refused);
 * convention 0 igraph, 1 NetworkX (refused), 2 sna, 3 UCINET. */
int32_t orc_centralization(const OGraph *g0, int32_t measure, int32_t mode, int32_t loops,
                           int32_t normalized, int32_t convention, double *out, char *err,
                           size_t err_len) {
    char msg[256];
    if (!g0) { write_err(err, err_len, "graph handle is null"); return ORC_FAILED; }
    if (measure < 0 || measure > 3) {
        snprintf(msg, sizeof msg, "unknown centralization measure code %d", measure);
        write_err(err, err_len, msg);
        return ORC_FAILED;
    }
    if (mode < 1 || mode > 3) {
        snprintf(msg, sizeof msg, "unknown degree mode code %d", mode);
        write_err(err, err_len, msg);
        return ORC_FAILED;
    }
    if (convention < 0 || convention > 3) {
        snprintf(msg, sizeof msg, "unknown convention code %d", convention);
        write_err(err, err_len, msg);
        return ORC_FAILED;
    }
    if (convention == 1) {
        write_err(err, err_len,
                  "NetworkX has no centralization index; use the igraph, sna, or UCINET "
                  "convention");
        return ORC_FAILED;
    }
    if (measure == 3) {
        static const char *names[] = {"igraph", "NetworkX", "sna", "UCINET"};
        if (convention != 0) {
            snprintf(msg, sizeof msg,
                     "eigenvector centralization is provided for the igraph convention only, "
                     "not %s", names[convention]);
            write_err(err, err_len, msg);
        } else if (g0->directed && mode != 3) {
            write_err(err, err_len,
                      "eigenvector centralization is provided with direction ignored (mode "
                      "\"all\") only");
        } else {
            write_err(err, err_len, "eigenvector centralization is not part of this library");
        }
        return ORC_FAILED;
    }
    const OGraph *g = g0;
    OGraph *simple = NULL;
    if (convention == 2 || convention == 3) {
        if (simplify_max(g0, 0, &simple, err, err_len)) return ORC_FAILED;
        g = simple;
    }
    uint64_t n = g->n;
    int directed = g->directed;
    int32_t emode = directed ? mode : 3;
    int eloops = loops && convention == 0;
    double *scores = malloc((n ? n : 1) * sizeof *scores);
    int32_t rc = ORC_OK;
    double result = NAN;
    if (measure == 0) {
        if (convention == 3 && directed && emode == 3) {
            write_err(err, err_len,
                      "UCINET reports degree centralization for in- and out-degree only");
            rc = ORC_FAILED;
        } else {
            degree_scores(g, emode, eloops, scores);
        }
    } else if (measure == 1) {
        rc = orc_betweenness(g, 0, 0, 0, 0, scores, err, err_len);
    } else {
        rc = orc_closeness(g, 0, emode, 0, 0, 1, convention == 2 ? 2 : 0, scores, err, err_len);
        if (rc == ORC_OK && convention == 3 && n >= 2) {
            /* UCINET: NaN unless every node reaches every other (the "zero"
             * rule gives 0 exactly to the nodes that do not) */
            double *zero = malloc(n * sizeof *zero);
            rc = orc_closeness(g, 0, emode, 0, 0, 1, 2, zero, err, err_len);
            int full = 1;
            for (uint64_t v = 0; rc == ORC_OK && v < n; v++)
                if (!(zero[v] > 0.0)) full = 0;
            free(zero);
            if (rc == ORC_OK && !full) {
                *out = NAN;
                free(scores);
                orc_graph_free(simple);
                return ORC_OK;
            }
        }
    }
    if (rc == ORC_OK)
        result = freeman_index(n, directed, scores, measure, emode, loops, normalized,
                               convention);
    if (rc == ORC_OK) *out = result;
    free(scores);
    orc_graph_free(simple);
    return rc;
}

/* betweenness_max, bipartite.rs:228-241 (Borgatti and Everett 1997). */
static double twomode_betweenness_max(uint64_t own, uint64_t other) {
    if (other == 0 || own == 0) return NAN;
    __int128 m = other, s = (own - 1) / other, t = (own - 1) % other;
    __int128 twice = m * m * (s + 1) * (s + 1) + m * (s + 1) * (2 * t - s - 1) -
                     t * (2 * s - t + 3);
    return (double)twice / 2.0;
}

/* net_twomode_centrality, crates/capi/src/twomode.rs:237-286 ->
 * twomode_degree / twomode_closeness / twomode_betweenness
 * (bipartite.rs:130-269). types: n codes, 0 first mode, 1 second. measure
 * 0 degree, 1 closeness (unreachable 0 ignore, 1 Wasserman-Faust), 2
 * betweenness. */
int32_t orc_twomode_centrality(const OGraph *g, const int32_t *types, int32_t measure,
                               int32_t normalized, int32_t unreachable, double *out, char *err,
                               size_t err_len) {
    char msg[320];
    if (!g) { write_err(err, err_len, "graph handle is null"); return ORC_FAILED; }
    if (!types) { write_err(err, err_len, "node types are missing"); return ORC_FAILED; }
    uint64_t n = g->n;
    for (uint64_t v = 0; v < n; v++)
        if (types[v] != 0 && types[v] != 1) {
            snprintf(msg, sizeof msg,
                     "node %llu has type code %d; use 0 for the first mode and 1 for the second",
                     (unsigned long long)v, types[v]);
            write_err(err, err_len, msg);
            return ORC_FAILED;
        }
    if (measure == 1) {
        if (unreachable != 0 && unreachable != 1) {
            snprintf(msg, sizeof msg, "unknown code %d for unreachable nodes", unreachable);
            write_err(err, err_len, msg);
            return ORC_FAILED;
        }
    } else if (measure != 0 && measure != 2) {
        snprintf(msg, sizeof msg, "unknown two-mode centrality code %d", measure);
        write_err(err, err_len, msg);
        return ORC_FAILED;
    }
    for (uint64_t e = 0; e < g->m; e++)
        if (types[g->src[e]] == types[g->tgt[e]]) {
            snprintf(msg, sizeof msg,
                     "edge %llu joins nodes %u and %u, which are in the same mode; in a two-mode "
                     "network every edge joins the two modes",
                     (unsigned long long)e, g->src[e], g->tgt[e]);
            write_err(err, err_len, msg);
            return ORC_FAILED;
        }
    uint64_t second = 0;
    for (uint64_t v = 0; v < n; v++) second += types[v] == 1;
    uint64_t first = n - second;
    if (measure == 1 && unreachable == 1 && !normalized) {
        write_err(err, err_len, "the Wasserman-Faust rule applies to normalized closeness only");
        return ORC_FAILED;
    }
    OGraph *a = NULL;
    if (simplify_max(g, 1, &a, err, err_len)) return ORC_FAILED;
    int32_t rc = ORC_OK;
    if (measure == 0) {
        degree_scores(a, 3, 0, out);
        if (normalized)
            for (uint64_t v = 0; v < n; v++) out[v] /= (double)(types[v] ? first : second);
    } else if (measure == 1) {
        int64_t *all = malloc((n ? n : 1) * sizeof *all);
        uint64_t *reached = malloc((n ? n : 1) * sizeof *reached);
        double *total = malloc((n ? n : 1) * sizeof *total), *h = malloc((n ? n : 1) * sizeof *h),
               *lg = malloc((n ? n : 1) * sizeof *lg);
        for (uint64_t v = 0; v < n; v++) all[v] = (int64_t)v;
        rc = orc_distance_profiles(a, 0, 0, all, (int64_t)n, 0, reached, total, h, lg, err,
                                   err_len);
        for (uint64_t v = 0; rc == ORC_OK && v < n; v++) {
            if (reached[v] == 0) {
                out[v] = unreachable == 1 ? 0.0 : NAN;
                continue;
            }
            double own = (double)(types[v] ? second : first);
            double other = (double)(types[v] ? first : second);
            double c = normalized ? (other + 2.0 * (own - 1.0)) / total[v] : 1.0 / total[v];
            if (unreachable == 1) c *= (double)reached[v] / (double)(n - 1);
            out[v] = c;
        }
        free(all); free(reached); free(total); free(h); free(lg);
    } else {
        rc = orc_betweenness(a, 0, 0, 0, 0, out, err, err_len);
        if (rc == ORC_OK && normalized) {
            double max_first = twomode_betweenness_max(first, second);
            double max_second = twomode_betweenness_max(second, first);
            for (uint64_t v = 0; v < n; v++) {
                double mx = types[v] ? max_second : max_first;
                out[v] = mx > 0.0 ? out[v] / mx : NAN;
            }
        }
    }
    orc_graph_free(a);
    return rc;
}
