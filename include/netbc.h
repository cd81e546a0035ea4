/*
 * netbc: exact Brandes betweenness on an AMD Instinct MI355X (gfx950).
 *
 * Drop-in C ABI for knappnet's graph handle and betweenness entry point
 * (crates/capi/include/netcapi.h), plus the device path's own entry
 * points. Conventions follow netcapi.h:1-25:
 * - Functions that take `err` return 0 on success and 1 on failure, after
 *   writing a NUL-terminated message into `err` (at most `err_len` bytes;
 *   `err` may be NULL or `err_len` 0). Accessors return -1 for NULL.
 * - Node indices are 0-based; output buffers are allocated by the caller.
 * - A handle belongs to the caller until released with net_graph_free,
 *   which accepts NULL.
 * - weights_as: 0 = distance, 1 = strength (path lengths 1/w), used only
 *   when weighted != 0. `threads` is accepted for compatibility; the
 *   device computes every source and the result never depends on it.
 * Device failures (no GPU, out of memory, launch errors) return 1 with a
 * message; nothing aborts the process.
 */
#ifndef NETBC_H
#define NETBC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* An immutable graph; release with net_graph_free. The device copy of its
 * CSR is built on first use and freed with the handle. */
typedef struct GraphHandle GraphHandle;

/* Replaces net_graph_new, netcapi.h:43-48 (crates/capi/src/lib.rs:133-164,
 * crates/engine/src/graph.rs:142-234): builds a graph from m edges (src,
 * tgt, optional weights) on n nodes; directed != 0 keeps arc directions. */
int32_t net_graph_new(int64_t n, const int64_t *src, const int64_t *tgt,
                      const double *weights, int64_t m, int32_t directed,
                      GraphHandle **out, char *err, size_t err_len);

/* Replaces net_graph_free, netcapi.h:50-51 (lib.rs:166-175). */
void net_graph_free(GraphHandle *h);

/* Replaces net_graph_node_count, netcapi.h:53-54 (lib.rs:177-184). */
int64_t net_graph_node_count(const GraphHandle *h);

/* Replaces net_graph_edge_count, netcapi.h:56-57
 * (crates/capi/src/descriptive.rs:83-85). */
int64_t net_graph_edge_count(const GraphHandle *h);

/* Replaces net_betweenness, netcapi.h:386-391 (lib.rs:214-244,
 * crates/engine/src/metrics/betweenness.rs:63-115). Shortest-path
 * betweenness of every node, endpoints excluded, undirected pairs counted
 * once; normalized != 0 divides by (n-1)(n-2) (directed) or
 * (n-1)(n-2)/2. `out` holds n doubles. */
int32_t net_betweenness(const GraphHandle *h, int32_t weighted,
                        int32_t weights_as, int32_t normalized,
                        int32_t threads, double *out, char *err,
                        size_t err_len);

/* Replaces net_edge_betweenness, netcapi.h:393-399 (crates/capi/src/
 * roles.rs:77-100, crates/engine/src/metrics/reach.rs:174-232).
 * Shortest-path betweenness of every edge, in input order: an arc u -> v
 * on a shortest s-t path carries sigma_su * sigma_vt / sigma_st of the
 * pair, parallel edges share it, loops carry nothing, undirected edges
 * count each pair once; normalized != 0 divides by n(n-1) (directed) or
 * n(n-1)/2 (NaN below two nodes). `out` holds m doubles. `threads` is
 * accepted and ignored. */
int32_t net_edge_betweenness(const GraphHandle *h, int32_t weighted,
                             int32_t weights_as, int32_t normalized,
                             int32_t threads, double *out, char *err,
                             size_t err_len);

/* Replaces net_betweenness_subset, netcapi.h:1957-1979 (crates/capi/src/
 * centrality.rs:89-144, crates/engine/src/metrics/path_centrality.rs:
 * 421-689). Betweenness for sources and targets: node values (n doubles)
 * and, when `edges` is not NULL, edge values (m doubles). Node lists: NULL
 * or a negative count means every node. rescale: 0 unordered pairs, 1
 * NetworkX's betweenness_centrality (a source list is a sample), 2
 * NetworkX's *_subset; parallel: 0 distinct paths, 1 collapsed (NetworkX);
 * edge_split: 0 proportional, 1 NetworkX's subset edge split (a second
 * search pass carries its edge dependencies); tie: 0 relative 1e-10, 1
 * exact. Values are within 1e-12 of the reference's. `threads` is accepted
 * and ignored. Real lengths whose longest arc exceeds about 2,047 times
 * the shortest support neither exact ties nor the NetworkX edge split on
 * the device path (the call fails with a message). */
int32_t net_betweenness_subset(const GraphHandle *h, const int64_t *sources,
                               int64_t n_sources, const int64_t *targets,
                               int64_t n_targets, int32_t weighted,
                               int32_t weights_as, int32_t endpoints,
                               int32_t normalized, int32_t rescale,
                               int32_t parallel, int32_t edge_split,
                               int32_t tie, int32_t threads, double *nodes,
                               double *edges, char *err, size_t err_len);

/* Replaces net_closeness, netcapi.h:401-406 (crates/capi/src/
 * topography.rs:155-199, crates/engine/src/metrics/closeness.rs:131-481).
 * Closeness (harmonic == 0) or harmonic centrality (harmonic != 0) of every
 * node from its distances to the others: mode 1 along out-arcs, 2 along
 * in-arcs, 3 ignoring direction; unreachable (closeness only): 0 ignore
 * unreachable nodes, 1 Wasserman-Faust (normalized only), 2 zero unless
 * every other node is reached, 3 unreachable lengths count as the longest
 * length plus one (unweighted only). normalized != 0 scales closeness by
 * the reached count and harmonic centrality by 1/(n-1). Breadth-first
 * values, and real lengths whose longest arc exceeds about 2,047 times the
 * shortest (the reference's Dijkstra per source), match the reference's
 * bit for bit; other weighted runs add each source's reciprocal (and real)
 * lengths in node order instead of settle order (within 1e-12 relative).
 * `out` holds n doubles; `threads` is accepted and ignored. */
int32_t net_closeness(const GraphHandle *h, int32_t harmonic, int32_t mode,
                      int32_t weighted, int32_t weights_as, int32_t normalized,
                      int32_t unreachable, int32_t threads, double *out,
                      char *err, size_t err_len);

/* Replaces net_centralization, netcapi.h:534-539 (crates/capi/src/
 * topography.rs:134-159, crates/engine/src/metrics/topography.rs:376-568).
 * Freeman centralization: n * max - sum of the scores, divided by the
 * tool's theoretical maximum when normalized != 0. measure 0 degree (mode
 * 1 out, 2 in, 3 all; loops counted when loops != 0, igraph only), 1
 * betweenness (unweighted), 2 closeness (normalized; unreachable nodes
 * ignored, sna: zero; UCINET: NaN unless every node reaches every other);
 * convention 0 igraph, 1 NetworkX (an error, as in the reference), 2 sna
 * and 3 UCINET (parallel edges collapsed, loops ignored). Measure 3
 * (eigenvector) passes the reference's checks and then fails: it is not
 * part of this library. Writes one double; `threads` is ignored. */
int32_t net_centralization(const GraphHandle *h, int32_t measure, int32_t mode,
                           int32_t loops, int32_t normalized,
                           int32_t convention, int32_t threads, double *out,
                           char *err, size_t err_len);

/* Replaces net_twomode_centrality, netcapi.h:753-760 (crates/capi/src/
 * twomode.rs:229-286, crates/engine/src/metrics/bipartite.rs:130-269).
 * Degree (0), closeness (1) or betweenness (2) of the affiliation graph
 * (direction ignored, parallel edges collapsed), normalized for two-mode
 * data when normalized != 0 (Borgatti and Everett 1997). types: n codes, 0
 * first mode, 1 second; every edge must join the modes. unreachable
 * (closeness only): 0 ignore, 1 Wasserman-Faust (normalized only). `out`
 * holds n doubles; `threads` is ignored. */
int32_t net_twomode_centrality(const GraphHandle *h, const int32_t *types,
                               int32_t measure, int32_t normalized,
                               int32_t unreachable, int32_t threads,
                               double *out, char *err, size_t err_len);

/* ---- device-path entry points --------------------------------------- */

/* Run options. Zero-initialised means: FP64 dependencies, automatic batch
 * count and direction. */
typedef struct netbc_options {
    int32_t precision; /* 0: FP64 dependencies; 1: FP32 dependencies with
                          Kahan-compensated sums (sigma and the cross-source
                          sum stay FP64) */
    int32_t slots;     /* 64-source batches per launch; 0 = tuned default */
    int32_t direction; /* 0 automatic, 1 top-down only, 2 bottom-up when
                          the frontier allows */
    int32_t exact;     /* 1: the reference's summation order everywhere, so
                          FP64 scores match it bit for bit. 0 (default):
                          rows longer than 1024 arcs are summed in 256-arc
                          chunks whose partial sums are added in chunk
                          order, and runs of 4096 or more sources batch
                          sources whose searches look alike together
                          (summing them in that order). Both are
                          deterministic and change only the association of
                          FP64 sums (scores within 1e-12 relative); levels
                          and path counts are bit-exact either way. */
    double alpha;      /* direction switch factor; 0 = tuned default */
    int32_t sweep_group; /* entries per dependency-sweep task at most (1 to
                            64); 0 = tuned default */
    int32_t sweep_block; /* threads per workgroup of the dependency sweep
                            (64 to 1024, a multiple of 64); 0 = tuned */
    int32_t step_block;  /* threads per workgroup of the bottom-up step and
                            the hub tasks (64 to 1024, a multiple of 64);
                            0 = tuned */
    int32_t hub_in;      /* in-degree above which a node's path counts are
                            gathered by hub tasks (1 to 2^30); 0 = tuned */
    int32_t hub_chunk;   /* in-arcs per hub task (64 to 65536, a multiple of
                            64); 0 = tuned */
    int32_t bucket_width; /* integer lengths: distances per bucket of the
                             bucketed search (1 to the shortest arc, so that
                             every bucket is settled at once; wider values
                             are cut to the shortest arc); 0 = tuned */
    int32_t reserved;    /* must be 0 */
} netbc_options;

/* What a run saw. */
typedef struct netbc_diagnostics {
    double max_sigma;        /* largest shortest-path count */
    int32_t sigma_overflow;  /* 1 if a count exceeded 2^53 (FP64 no longer
                                exact; bit-exactness not guaranteed) */
    int32_t path_kind;       /* 0 breadth-first, 1 integer (Dial), 2 real */
    int64_t reached_nodes;   /* sum over sources of nodes reached */
    int64_t reached_arcs;    /* sum over sources of out-arcs of those nodes */
    int64_t levels;          /* search levels run */
    double kernel_ms;        /* device time of the search and reduction */
    double dominant_ms;      /* device time of the dependency sweep
                                (closeness: of the bottom-up steps) */
    int64_t dominant_launches;
    int32_t slots;           /* 64-source batches per launch used */
    int32_t bottom_up_levels; /* search levels run bottom-up */
    double build_ms;         /* device CSR/CSC build time (first call) */
    int64_t work[16];        /* engine work counters; zero unless the library
                                is built with NETBC_COUNTERS */
    int64_t dominant_bytes;  /* closeness runs: algorithmic bytes of the
                                timed kernel (the bottom-up step of the
                                breadth-first search), summed over
                                launches */
} netbc_diagnostics;

/* Betweenness restricted to a source list (NULL with k = n: every node,
 * identical to net_betweenness). The per-source searches and the ordered
 * group sum (crates/engine/src/parallel.rs:383-503) run over list
 * positions; scaling is betweenness.rs:100-113's. Same error conventions;
 * `diag` may be NULL. No counterpart in netcapi.h: it is the device path's
 * extended entry (sampled sources, precision choice, diagnostics). */
int32_t netbc_betweenness_sources(const GraphHandle *h, const int64_t *sources,
                                  int64_t k, int32_t weighted,
                                  int32_t weights_as, int32_t normalized,
                                  const netbc_options *opt, double *out,
                                  netbc_diagnostics *diag, char *err,
                                  size_t err_len);

/* Edge betweenness restricted to a source list (NULL: every node,
 * identical to net_edge_betweenness), scaled as reach.rs:207-224; `out`
 * holds m doubles. Options and diagnostics as netbc_betweenness_sources
 * (`exact` orders each node's dependency sum as the reference; the edge
 * sums are always within 1e-12 of the reference's, not bit-identical). No
 * counterpart in netcapi.h. */
int32_t netbc_edge_betweenness_sources(const GraphHandle *h,
                                       const int64_t *sources, int64_t k,
                                       int32_t weighted, int32_t weights_as,
                                       int32_t normalized,
                                       const netbc_options *opt, double *out,
                                       netbc_diagnostics *diag, char *err,
                                       size_t err_len);

/* net_betweenness_subset with run options and diagnostics (both may be
 * NULL). No counterpart in netcapi.h. */
int32_t netbc_betweenness_subset_ex(const GraphHandle *h,
                                    const int64_t *sources, int64_t n_sources,
                                    const int64_t *targets, int64_t n_targets,
                                    int32_t weighted, int32_t weights_as,
                                    int32_t endpoints, int32_t normalized,
                                    int32_t rescale, int32_t parallel,
                                    int32_t edge_split, int32_t tie,
                                    const netbc_options *opt, double *nodes,
                                    double *edges, netbc_diagnostics *diag,
                                    char *err, size_t err_len);

/* net_closeness with run options and diagnostics (both may be NULL). No
 * counterpart in netcapi.h. */
int32_t netbc_closeness_ex(const GraphHandle *h, int32_t harmonic,
                           int32_t mode, int32_t weighted, int32_t weights_as,
                           int32_t normalized, int32_t unreachable,
                           const netbc_options *opt, double *out,
                           netbc_diagnostics *diag, char *err, size_t err_len);

/* Betweenness of many independent graphs in one call (null-model and
 * permutation replicates). Graph i has node_counts[i] nodes and
 * edge_counts[i] edges, listed one graph after another in src/tgt (node
 * ids local to the graph) and weights (NULL, or one per edge). `out`
 * receives the graphs' scores one after another (sum of node_counts
 * doubles), each exactly what net_betweenness returns for that graph
 * alone. Unweighted graphs of up to 2048 nodes run together in one
 * persistent launch; others run one at a time. Errors name the graph
 * ("graph 3: edge 5 refers to ..."). No counterpart in netcapi.h: it
 * replaces the one-call-per-replicate loop of r/R/cug.R:42-47. */
int32_t netbc_betweenness_batch(int64_t graphs, const int64_t *node_counts,
                                const int64_t *edge_counts, const int64_t *src,
                                const int64_t *tgt, const double *weights,
                                int32_t directed, int32_t weighted,
                                int32_t weights_as, int32_t normalized,
                                const netbc_options *opt, double *out,
                                netbc_diagnostics *diag, char *err,
                                size_t err_len);

/* A plan for repeated netbc_betweenness_batch calls on one batch shape:
 * the same graph count, node counts and edge counts every time, as the
 * replicate batches of a conditional uniform graph test have (r/R/cug.R:
 * 25-47 draws `reps` graphs of the observed size and edge count per call).
 * netbc_batch_plan_new allocates every device buffer the shape needs and,
 * with capture != 0, records the whole device side of a call (building
 * the union graph's CSR from its edges, the persistent launch over the
 * small graphs and the ordered reduction) into a HIP graph; with
 * capture == 0 the same kernels are launched one by one on every run.
 * Unweighted graphs only; `opt` may be NULL. No counterpart in netcapi.h. */
typedef struct netbc_batch_plan netbc_batch_plan;
int32_t netbc_batch_plan_new(int64_t graphs, const int64_t *node_counts,
                             const int64_t *edge_counts, int32_t directed,
                             int32_t normalized, int32_t capture,
                             const netbc_options *opt, netbc_batch_plan **plan,
                             char *err, size_t err_len);

/* One batch through a plan: src/tgt hold the edges of every graph in the
 * layout netbc_betweenness_batch takes (the plan's edge counts, node ids
 * local to each graph); `out` receives the scores, equal bit for bit to
 * netbc_betweenness_batch's on the same graphs (weighted = 0). Graphs the
 * persistent launch does not take (more than 2048 nodes, searches past 254
 * levels) are computed one at a time after the replay. Calls on one plan
 * are serialised; `diag` may be NULL. */
int32_t netbc_batch_plan_run(netbc_batch_plan *plan, const int64_t *src,
                             const int64_t *tgt, double *out,
                             netbc_diagnostics *diag, char *err,
                             size_t err_len);

/* Releases a plan (NULL is a no-op). */
void netbc_batch_plan_free(netbc_batch_plan *plan);

/* netbc_betweenness_batch without device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The graphs arrive already on the device as their union graph
 * in the layout net_graph_new builds and netbc_graph_csr returns: graph i
 * has node_counts[i] nodes (host array), numbered after those of graphs
 * 0..i-1; `offsets` (total nodes + 1) and `targets` (`arcs` entries) are the
 * union's out-CSR, `in_offsets` / `in_targets` its in-CSR (directed only;
 * NULL when undirected); `out` receives the scores (total nodes doubles) on
 * the device, equal bit for bit to what netbc_betweenness_batch returns for
 * the same unweighted graphs. Limits: at most 65,536 graphs per call, each
 * of at most 2048 nodes and searched within 254 levels (the call fails
 * with a message otherwise); rows and offsets are checked on the device
 * (an arc that leaves its graph is an error). Scratch: one static 64 MiB
 * device array per GPU the library runs on, held until the process ends;
 * every call on a GPU shares it, so calls that overlap on one GPU race
 * (serialise them). Runs on `stream` (NULL: the default stream) and waits
 * for it before returning. */
int32_t netbc_betweenness_batch_noalloc(int64_t graphs,
                                        const int64_t *node_counts,
                                        int32_t directed, int32_t normalized,
                                        const uint32_t *offsets,
                                        const uint32_t *targets,
                                        const uint32_t *in_offsets,
                                        const uint32_t *in_targets,
                                        int64_t arcs, double *out, void *stream,
                                        char *err, size_t err_len);

/* Debug entry: the per-source state of k sources, each row n long:
 * dist = hop count (breadth-first) or integer distance (Dial),
 * UINT64_MAX when unreached; sigma = path counts; delta = dependencies
 * (0 for unreached nodes and for the source itself). Follows the
 * reference's Workspace state (betweenness.rs:161-295); not for real
 * lengths. */
int32_t netbc_state_dump(const GraphHandle *h, int32_t weighted,
                         int32_t weights_as, int32_t precision,
                         const int64_t *sources, int64_t k, uint64_t *dist,
                         double *sigma, double *delta, char *err,
                         size_t err_len);

/* Number of stored arcs (m directed; 2m - loops undirected), or -1. */
int64_t netbc_graph_arc_count(const GraphHandle *h);

/* Copies the device CSR (incoming == 0) or CSC (incoming != 0) back:
 * offsets n+1, targets and (if the graph is weighted and weights is not
 * NULL) weights, one per arc. Mirrors crates/engine/src/graph.rs:18-119's
 * Csr layout, for equality tests against the host build. */
int32_t netbc_graph_csr(const GraphHandle *h, int32_t incoming,
                        uint32_t *offsets, uint32_t *targets, double *weights,
                        char *err, size_t err_len);

/* Diagnostics: copies `bytes` (a multiple of 16) from one device buffer to
 * another with a plain streaming kernel (netbc_copy_kernel), so hardware
 * traffic counters read in the same profiling run can be scaled to bytes.
 * Both pointers must be device allocations. No counterpart in netcapi.h. */
int32_t netbc_calibration_copy(void *dst, const void *src, size_t bytes,
                               char *err, size_t err_len);

/* Diagnostics: reads one 8-byte word from each `line`-byte line of `src`
 * (`bytes` / `line` lines, a power of two, every line once, in a scattered
 * order) with netbc_gather_kernel, writing one 8-byte sum per thread
 * (524,288 threads: `dst` holds 4 MiB). Counters read for it say how many
 * bytes a random 8-byte read costs. No counterpart in netcapi.h. */
int32_t netbc_calibration_gather(void *dst, const void *src, size_t bytes,
                                 size_t line, char *err, size_t err_len);

/* Name of the device the library runs on (e.g. "AMD Instinct MI355X,
 * gfx950, 256 CUs"). */
int32_t netbc_device_info(char *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* NETBC_H */
