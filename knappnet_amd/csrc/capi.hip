// The C ABI: graph handles and the betweenness entry points.
//
// Behaviour follows crates/capi/src/lib.rs: status 0 / 1, messages
// truncated to err_len - 1 bytes plus NUL, NULL err tolerated (60-71),
// failures caught and reported instead of unwinding (73-85), the same
// validation order and texts as Graph::from_edges (graph.rs:142-234) and
// error.rs:62-99.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>

#include "../../include/netbc.h"
#include "netbc_internal.h"
#include "small.h"

#define NETBC_EXPORT extern "C" __attribute__((visibility("default")))

struct GraphHandle {
    netbc::Graph g;
};

namespace {

constexpr int32_t kOk = 0;
constexpr int32_t kFailed = 1;

void write_err(char* err, size_t err_len, const std::string& msg) {
    if (!err || err_len == 0) return;
    size_t n = std::min(msg.size(), err_len - 1);
    memcpy(err, msg.data(), n);
    err[n] = 0;
}

template <class F>
int32_t guard(char* err, size_t err_len, F&& f) {
    try {
        f();
        return kOk;
    } catch (const netbc::Error& e) {
        write_err(err, err_len, e.what());
    } catch (const std::bad_alloc&) {
        write_err(err, err_len, "out of host memory");
    } catch (...) {
        write_err(err, err_len, "internal error (panic) in the network engine");
    }
    return kFailed;
}

netbc::Graph& handle(const GraphHandle* h) {
    if (!h) throw netbc::Error("graph handle is null");
    return const_cast<GraphHandle*>(h)->g;
}

void check_kind(int32_t weights_as) {
    if (weights_as != 0 && weights_as != 1)
        throw netbc::Error("unknown weight kind code " + std::to_string(weights_as));
}

// Selects a device, failing with a message when there is none.
void ensure_device() {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        throw netbc::Error("no HIP device is available for the betweenness engine");
}

double prepare(netbc::Graph& g) {
    ensure_device();
    auto t0 = std::chrono::steady_clock::now();
    bool had = g.dev_ready;
    netbc::build_device_graph(g);
    if (had) return 0.0;
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

void fill_diag(netbc_diagnostics* out, const netbc::RunDiag& d, double build_ms) {
    if (!out) return;
    out->max_sigma = d.max_sigma;
    out->sigma_overflow = d.sigma_overflow;
    out->path_kind = d.path_kind;
    out->reached_nodes = d.reached_nodes;
    out->reached_arcs = d.reached_arcs;
    out->levels = d.levels;
    out->kernel_ms = d.kernel_ms;
    out->dominant_ms = d.dominant_ms;
    out->dominant_launches = d.dominant_launches;
    out->slots = d.slots;
    out->bottom_up_levels = (int32_t)d.bottom_up_levels;
    out->build_ms = build_ms;
    for (int i = 0; i < 16; i++) out->work[i] = d.work[i];
    out->dominant_bytes = d.dominant_bytes;
}

// Rejects option values the engine does not define.
void check_options(const netbc_options* opt) {
    if (!opt) return;
    if (opt->precision < 0 || opt->precision > 1)
        throw netbc::Error("unknown precision code " + std::to_string(opt->precision));
    if (opt->slots < 0) throw netbc::Error("negative slot count");
    if (opt->direction < 0 || opt->direction > 2)
        throw netbc::Error("unknown direction code " + std::to_string(opt->direction));
    if (opt->sweep_group < 0 || opt->sweep_group > 64)
        throw netbc::Error("sweep group " + std::to_string(opt->sweep_group) +
                           " is outside 0..64");
    if (!(opt->alpha >= 0.0)) throw netbc::Error("negative or NaN direction factor");
    for (const auto& [v, what] : {std::pair<int32_t, const char*>{opt->sweep_block, "sweep"},
                                  {opt->step_block, "step"}})
        if (v < 0 || v > 1024 || v % 64)
            throw netbc::Error(std::string(what) + " block " + std::to_string(v) +
                               " is not 0 or a multiple of 64 up to 1024");
    if (opt->hub_in < 0 || opt->hub_in > (1 << 30))
        throw netbc::Error("hub in-degree " + std::to_string(opt->hub_in) +
                           " is outside 0..2^30");
    if (opt->hub_chunk < 0 || opt->hub_chunk > 65536 || opt->hub_chunk % 64)
        throw netbc::Error("hub chunk " + std::to_string(opt->hub_chunk) +
                           " is not 0 or a multiple of 64 up to 65536");
    if (opt->bucket_width < 0) throw netbc::Error("negative bucket width");
    if (opt->reserved != 0) throw netbc::Error("the reserved option field must be 0");
}

netbc::RunOptions options(const netbc_options* opt) {
    netbc::RunOptions o;
    if (opt) {
        o.precision = opt->precision;
        o.slots = opt->slots;
        o.direction = opt->direction;
        o.alpha = opt->alpha;
        o.exact = opt->exact;
        o.sweep_group = opt->sweep_group;
        o.sweep_block = opt->sweep_block;
        o.step_block = opt->step_block;
        o.hub_in = opt->hub_in;
        o.hub_chunk = opt->hub_chunk;
        o.bucket_width = opt->bucket_width;
    }
    return o;
}

void merge_diag(netbc::RunDiag& into, const netbc::RunDiag& d) {
    into.max_sigma = std::max(into.max_sigma, d.max_sigma);
    into.sigma_overflow |= d.sigma_overflow;
    into.path_kind = std::max(into.path_kind, d.path_kind);
    into.reached_nodes += d.reached_nodes;
    into.reached_arcs += d.reached_arcs;
    into.levels += d.levels;
    into.bottom_up_levels += d.bottom_up_levels;
    into.kernel_ms += d.kernel_ms;
    into.dominant_ms += d.dominant_ms;
    into.dominant_launches += d.dominant_launches;
    into.dominant_bytes += d.dominant_bytes;
    into.slots = std::max(into.slots, d.slots);
    for (int i = 0; i < 16; i++) into.work[i] += d.work[i];
}

// Betweenness of one handle over a source list (sources == nullptr: item i
// is node i), through the batched engine; returns the run's diagnostics.
// With edge, out receives the m edge scores instead of the n node scores.
netbc::RunDiag run_graph(netbc::Graph& g, const int64_t* sources, int64_t k, int32_t weighted,
                         int32_t weights_as, int32_t normalized, const netbc_options* opt,
                         double* out, double* build_ms, bool edge = false) {
    std::lock_guard<std::mutex> lock(g.mu);
    if (weighted && !g.has_weights) throw netbc::Error("the graph has no edge weights");
    double b = prepare(g);
    if (build_ms) *build_ms = b;
    netbc::Lengths L;
    netbc::classify_lengths(g, weighted != 0, weights_as, L);
    netbc::RunDiag d;
    netbc::DevBuf<double> bc, eb;
    bc.alloc(g.n);
    if (edge) eb.alloc(g.m);
    netbc::run_batched(g, L, sources, k, options(opt), bc.p, d, edge ? eb.p : nullptr);
    if (edge) netbc::finish_edge_scores(g, normalized != 0, eb.p, out);
    else netbc::finish_scores(g, normalized != 0, bc.p, out);
    return d;
}

void run_sources(netbc::Graph& g, const int64_t* sources, int64_t k, int32_t weighted,
                 int32_t weights_as, int32_t normalized, const netbc_options* opt, double* out,
                 netbc_diagnostics* diag, bool edge = false) {
    double build_ms = 0.0;
    netbc::RunDiag d = run_graph(g, sources, k, weighted, weights_as, normalized, opt, out,
                                 &build_ms, edge);
    fill_diag(diag, d, build_ms);
}

// Checks a caller's source list; returns the item count (all n nodes
// when sources is null).
int64_t check_sources(const netbc::Graph& g, const int64_t* sources, int64_t k) {
    if (!sources) return (int64_t)g.n;
    if (k < 0) throw netbc::Error("source count must be non-negative");
    for (int64_t i = 0; i < k; i++)
        if (sources[i] < 0 || (uint64_t)sources[i] >= g.n)
            throw netbc::Error("source " + std::to_string(i) + " is node index " +
                               std::to_string(sources[i]) + ", but valid indices are 0.." +
                               std::to_string(g.n));
    return k;
}

// The reference's scale for one graph (betweenness.rs:100-113).
double graph_scale(uint64_t n, bool directed, bool normalized) {
    double scale = directed ? 1.0 : 0.5;
    if (normalized && n > 2) {
        double pairs = (double)((n - 1) * (n - 2));
        scale *= directed ? 1.0 / pairs : 2.0 / pairs;
    }
    return scale;
}

// ---- betweenness for sources and targets ----------------------------------

// A node list of net_betweenness_subset (centrality.rs:45-63): NULL or a
// negative count means every node.
void check_node_list(const netbc::Graph& g, const int64_t* nodes, int64_t k) {
    if (!nodes || k < 0) return;
    for (int64_t i = 0; i < k; i++)
        if (nodes[i] < 0 || (uint64_t)nodes[i] >= g.n)
            throw netbc::Error("node index " + std::to_string(nodes[i]) +
                               " is out of range for a network with " + std::to_string(g.n) +
                               " nodes");
}

// Edge lengths of the path graph (paths/mod.rs:175-192): 1, w or 1/w, with
// PathGraph::new's strength check and require_positive's (paths/mod.rs:
// 416-425), which names the first arc in row order whose length is not
// positive (with collapse, the arc kept for its adjacent pair).
std::vector<double> subset_lengths(const netbc::Graph& g, bool weighted, int weights_as,
                                   bool collapse) {
    const uint64_t m = g.m;
    std::vector<double> len(m, 1.0);
    if (!weighted) return len;
    if (!g.has_weights) throw netbc::Error("the graph has no edge weights");
    bool bad = false;
    for (uint64_t e = 0; e < m; e++) {
        const double w = g.w[e];
        if (weights_as == 1) {
            if (w <= 0.0)
                throw netbc::Error("edge " + std::to_string(e) + " has weight " +
                                   netbc::format_f64(w) +
                                   "; weights read as strengths become lengths 1/w and must be "
                                   "strictly positive");
            len[e] = 1.0 / w;
        } else {
            len[e] = w;
        }
        bad |= !(len[e] > 0.0);
    }
    if (!bad) return len;
    // arcs in row order: (tail, head, generation), generation = edge order,
    // an undirected edge's reverse right after it
    struct Arc { uint32_t t, h; uint64_t gen; uint32_t e; };
    std::vector<Arc> arcs;
    for (uint64_t e = 0; e < m; e++) {
        arcs.push_back({g.src[e], g.tgt[e], arcs.size(), (uint32_t)e});
        if (!g.directed && g.src[e] != g.tgt[e])
            arcs.push_back({g.tgt[e], g.src[e], arcs.size(), (uint32_t)e});
    }
    std::sort(arcs.begin(), arcs.end(), [](const Arc& a, const Arc& b) {
        return a.t != b.t ? a.t < b.t : a.h != b.h ? a.h < b.h : a.gen < b.gen;
    });
    for (size_t i = 0; i < arcs.size();) {
        size_t j = i + 1;
        if (collapse)
            while (j < arcs.size() && arcs[j].t == arcs[i].t && arcs[j].h == arcs[i].h) j++;
        uint32_t keep = arcs[i].e;  // the first arc of least length
        for (size_t q = i + 1; q < j; q++)
            if (len[arcs[q].e] < len[keep]) keep = arcs[q].e;
        if (!(len[keep] > 0.0))
            throw netbc::Error("edge " + std::to_string(keep) + " has length " +
                               netbc::format_f64(len[keep]) +
                               "; shortest-path betweenness needs strictly positive lengths");
        i = j;
    }
    return len;
}

// The path graph with parallel arcs collapsed (paths/mod.rs:203-215): one
// edge per adjacent pair (unordered in undirected graphs), the first of
// least length; kept[i] is the input edge behind edge i.
void collapse_graph(const netbc::Graph& g, const std::vector<double>& len, netbc::Graph& out,
                    std::vector<uint32_t>& kept) {
    const uint64_t m = g.m;
    std::vector<std::pair<uint64_t, uint32_t>> key(m);
    for (uint64_t e = 0; e < m; e++) {
        uint64_t a = g.src[e], b = g.tgt[e];
        if (!g.directed && a > b) std::swap(a, b);
        key[e] = {a << 32 | b, (uint32_t)e};
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const auto& x, const auto& y) { return x.first < y.first; });
    kept.clear();
    for (size_t i = 0; i < m;) {
        size_t j = i;
        uint32_t best = key[i].second;
        while (j < m && key[j].first == key[i].first) {
            if (len[key[j].second] < len[best]) best = key[j].second;
            j++;
        }
        kept.push_back(best);
        i = j;
    }
    std::sort(kept.begin(), kept.end());
    out.n = g.n;
    out.m = kept.size();
    out.directed = g.directed;
    out.has_weights = g.has_weights;
    out.src.resize(out.m);
    out.tgt.resize(out.m);
    if (g.has_weights) out.w.resize(out.m);
    for (size_t i = 0; i < kept.size(); i++) {
        out.src[i] = g.src[kept[i]];
        out.tgt[i] = g.tgt[kept[i]];
        if (g.has_weights) out.w[i] = g.w[kept[i]];
    }
}

// NetworkX's _rescale (path_centrality.rs:596-649) as factors: values of
// sources get f_src, others f_other (equal unless a sample is rescaled);
// sampled < 0: no sample.
void nx_factors(uint64_t n, bool normalized, bool directed, bool endpoints, int64_t sampled,
                double& f_src, double& f_other) {
    f_src = f_other = 1.0;
    const double big_n = endpoints ? (double)n : (double)n - 1.0;
    if (big_n < 2.0) return;
    const double correction = directed ? 1.0 : 2.0;
    const double k = sampled >= 0 ? (double)sampled : big_n;
    if (sampled < 0 || endpoints) {
        f_src = f_other = normalized ? 1.0 / (k * (big_n - 1.0)) : big_n / (k * correction);
        return;
    }
    if (normalized) {
        f_src = k > 1.0 ? 1.0 / ((k - 1.0) * (big_n - 1.0)) : NAN;
        f_other = 1.0 / (k * (big_n - 1.0));
    } else {
        f_src = k > 1.0 ? big_n / ((k - 1.0) * correction) : NAN;
        f_other = big_n / (k * correction);
    }
}

// betweenness_subset, path_centrality.rs:421-593, through the batched
// engine: searches, target-aware dependencies, endpoint counts, the
// cross-source sums and the scaling run on the device; the collapsed
// path graph is built on the host, and edge values of collapsed pairs
// are shared among their parallel edges there (path_centrality.rs:651-689).
netbc::RunDiag run_subset(netbc::Graph& g, const int64_t* sources, int64_t n_sources,
                          const int64_t* targets, int64_t n_targets, int32_t weighted,
                          int32_t weights_as, int32_t endpoints, int32_t normalized,
                          int32_t rescale, int32_t parallel, int32_t edge_split, int32_t tie,
                          const netbc_options* opt, double* nodes, double* edges) {
    std::lock_guard<std::mutex> lock(g.mu);
    const bool collapse = parallel == 1;
    const std::vector<double> len = subset_lengths(g, weighted != 0, weights_as, collapse);
    const uint64_t n = g.n;
    netbc::Graph derived;
    std::vector<uint32_t> kept;
    netbc::Graph* pg = &g;
    if (collapse) {
        collapse_graph(g, len, derived, kept);
        pg = &derived;
    }
    prepare(*pg);
    netbc::Lengths L;
    netbc::classify_lengths(*pg, weighted != 0, weights_as, L);
    const bool all_src = !sources || n_sources < 0;
    const int64_t k = all_src ? (int64_t)n : n_sources;
    netbc::DevBuf<uint8_t> tflag, sflag;
    netbc::SubsetRun sub;
    sub.endpoints = endpoints != 0;
    sub.tie_exact = tie == 1;
    if (targets && n_targets >= 0) {
        std::vector<uint8_t> f(std::max<uint64_t>(n, 1), 0);
        for (int64_t i = 0; i < n_targets; i++) f[targets[i]] = 1;
        tflag.alloc(f.size());
        NETBC_HIP(hipMemcpyAsync(tflag.p, f.data(), f.size(), hipMemcpyHostToDevice, pg->stream));
        sub.d_target = tflag.p;
    }
    netbc::DevBuf<double> bc, eb;
    bc.alloc(n);
    if (edges) eb.alloc(pg->m);
    netbc::RunDiag d;
    const bool quirk = edges && edge_split == 1;
    netbc::run_batched(*pg, L, all_src ? nullptr : sources, k, options(opt), bc.p, d,
                       edges && !quirk ? eb.p : nullptr, &sub);
    if (quirk) {
        // NetworkX's edge split: a second run whose sweep carries the edge
        // dependencies instead (its node sums are not betweenness)
        netbc::SubsetRun sq = sub;
        sq.endpoints = false;
        sq.edge_split_nx = true;
        netbc::DevBuf<double> scratch;
        scratch.alloc(n);
        netbc::RunDiag d2;
        netbc::run_batched(*pg, L, all_src ? nullptr : sources, k, options(opt), scratch.p, d2,
                           eb.p, &sq);
    }
    // scaling, path_centrality.rs:527-574
    double fn_src, fn_other, fe, fe_unused;
    const bool directed = g.directed;
    if (rescale == 0) {
        const double half = directed ? 1.0 : 0.5;
        double np = 1.0, ep = 1.0;
        if (normalized) {
            const double nf = (double)n;
            np = endpoints ? nf * (nf - 1.0) : (nf - 1.0) * (nf - 2.0);
            ep = nf * (nf - 1.0);
        }
        fn_src = fn_other = np > 0.0 ? half / np : half;
        fe = ep > 0.0 ? half / ep : half;
    } else {
        const int64_t sampled = rescale == 1 && !all_src ? k : -1;
        nx_factors(n, normalized != 0, directed, rescale == 1 && endpoints, sampled, fn_src,
                   fn_other);
        nx_factors(n, normalized != 0, directed, true, sampled, fe, fe_unused);
    }
    if (n) {
        if (fn_src != fn_other) {
            std::vector<uint8_t> f(n, 0);
            for (int64_t i = 0; i < k; i++) f[all_src ? i : sources[i]] = 1;
            sflag.alloc(n);
            NETBC_HIP(hipMemcpyAsync(sflag.p, f.data(), n, hipMemcpyHostToDevice, pg->stream));
        }
        netbc::scale_flagged(*pg, bc.p, n, sflag.p, fn_src, fn_other);
        NETBC_HIP(hipMemcpyAsync(nodes, bc.p, n * sizeof(double), hipMemcpyDeviceToHost,
                                 pg->stream));
    }
    std::vector<double> ev(edges ? pg->m : 0);
    if (edges && pg->m) {
        netbc::scale_flagged(*pg, eb.p, pg->m, nullptr, fe, fe);
        NETBC_HIP(hipMemcpyAsync(ev.data(), eb.p, pg->m * sizeof(double), hipMemcpyDeviceToHost,
                                 pg->stream));
    }
    NETBC_HIP(hipStreamSynchronize(pg->stream));
    if (!edges) return d;
    if (!collapse) {
        std::copy(ev.begin(), ev.end(), edges);
        return d;
    }
    // share_among_parallel: a pair's value equally among its parallel edges
    // of least length; other parallels and loops get 0
    std::vector<double> vals(g.m, 0.0);
    for (size_t i = 0; i < kept.size(); i++) vals[kept[i]] = ev[i];
    std::vector<std::pair<uint64_t, uint32_t>> key(g.m);
    for (uint64_t e = 0; e < g.m; e++) {
        uint64_t a = g.src[e], b = g.tgt[e];
        if (!directed && a > b) std::swap(a, b);
        key[e] = {a << 32 | b, (uint32_t)e};
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const auto& x, const auto& y) { return x.first < y.first; });
    std::fill(edges, edges + g.m, 0.0);
    for (size_t i = 0; i < g.m;) {
        size_t j = i;
        while (j < g.m && key[j].first == key[i].first) j++;
        const uint64_t a = key[i].first >> 32, b = key[i].first & 0xFFFFFFFFull;
        if (a != b) {
            double total = 0.0, shortest = INFINITY;
            for (size_t q = i; q < j; q++) total += vals[key[q].second];
            for (size_t q = i; q < j; q++) shortest = std::min(shortest, len[key[q].second]);
            double ties = 0.0;
            for (size_t q = i; q < j; q++) ties += len[key[q].second] == shortest;
            for (size_t q = i; q < j; q++)
                if (len[key[q].second] == shortest) edges[key[q].second] = total / ties;
        }
        i = j;
    }
    return d;
}

// The checks of net_betweenness_subset in the reference's order
// (centrality.rs:108-136), then the run.
void subset_entry(const GraphHandle* h, const int64_t* sources, int64_t n_sources,
                  const int64_t* targets, int64_t n_targets, int32_t weighted, int32_t weights_as,
                  int32_t endpoints, int32_t normalized, int32_t rescale, int32_t parallel,
                  int32_t edge_split, int32_t tie, const netbc_options* opt, double* nodes,
                  double* edges, netbc_diagnostics* diag) {
    netbc::Graph& g = handle(h);
    check_node_list(g, sources, n_sources);
    check_node_list(g, targets, n_targets);
    check_kind(weights_as);
    if (rescale < 0 || rescale > 2)
        throw netbc::Error("unknown rescaling code " + std::to_string(rescale));
    if (parallel < 0 || parallel > 1)
        throw netbc::Error("unknown code " + std::to_string(parallel) + " for parallel edges");
    if (edge_split < 0 || edge_split > 1)
        throw netbc::Error("unknown edge split code " + std::to_string(edge_split));
    if (tie < 0 || tie > 1) throw netbc::Error("unknown tie code " + std::to_string(tie));
    check_options(opt);
    if (weighted && !g.has_weights) throw netbc::Error("the graph has no edge weights");
    netbc::RunDiag d = run_subset(g, sources, n_sources, targets, n_targets, weighted,
                                  weights_as, endpoints, normalized, rescale, parallel,
                                  edge_split, tie, opt, nodes, edges);
    fill_diag(diag, d, 0.0);
}

// Profiles and values of a closeness run (no option checks): the graph
// oriented for the mode (closeness.rs:225-238: in-paths are the out-paths
// of the reversed graph, paths ignoring direction those of the undirected
// graph), the path lengths, then the device run. Returns the device build
// time of the graph searched.
double closeness_run(netbc::Graph& g, int32_t mode, int32_t weighted, int32_t weights_as,
                     netbc::ClosenessRun cr, const netbc_options* opt, double* out,
                     netbc::RunDiag& d) {
    std::lock_guard<std::mutex> lock(g.mu);
    netbc::Graph derived;
    netbc::Graph* pg = &g;
    if (g.directed && mode != 1) {
        derived.n = g.n;
        derived.m = g.m;
        derived.directed = mode == 2;
        derived.has_weights = g.has_weights;
        derived.src = mode == 2 ? g.tgt : g.src;
        derived.tgt = mode == 2 ? g.src : g.tgt;
        derived.w = g.w;
        pg = &derived;
    }
    const double build_ms = prepare(*pg);
    netbc::Lengths L;
    netbc::classify_lengths(*pg, weighted != 0, weights_as, L);
    // uniform weights: hop counts times the common length
    // (closeness.rs:325-344)
    if (L.kind == netbc::kUniform && weighted && pg->m > 0)
        cr.unit = weights_as == 1 ? 1.0 / pg->w[0] : pg->w[0];
    netbc::run_closeness(*pg, L, options(opt), cr, out, d);
    return build_ms;
}

// simplify, topography.rs:105-140: loops dropped, one edge per ordered
// pair (unordered, smaller id first, when `undirected` or the graph is),
// the largest weight of a pair kept, edges sorted by (source, target).
void simplify_graph(const netbc::Graph& g, bool undirected, netbc::Graph& out) {
    const bool directed = g.directed && !undirected;
    struct Arc {
        uint32_t s, t;
        double x;
    };
    std::vector<Arc> a;
    a.reserve(g.m);
    for (uint64_t e = 0; e < g.m; e++) {
        uint32_t s = g.src[e], t = g.tgt[e];
        if (s == t) continue;
        const double x = g.has_weights ? g.w[e] : 1.0;
        if (!directed && s > t) std::swap(s, t);
        a.push_back({s, t, x});
    }
    std::sort(a.begin(), a.end(),
              [](const Arc& p, const Arc& q) { return p.s != q.s ? p.s < q.s : p.t < q.t; });
    out.n = g.n;
    out.directed = directed;
    out.has_weights = g.has_weights;
    out.src.clear();
    out.tgt.clear();
    out.w.clear();
    for (const Arc& r : a) {
        if (!out.src.empty() && out.src.back() == r.s && out.tgt.back() == r.t) {
            if (g.has_weights) out.w.back() = std::fmax(out.w.back(), r.x);
            continue;
        }
        out.src.push_back(r.s);
        out.tgt.push_back(r.t);
        if (g.has_weights) out.w.push_back(r.x);
    }
    out.m = out.src.size();
}

// centralization_max, topography.rs:504-568 (measure 0 degree, 1
// betweenness, 2 closeness; convention 0 igraph, 2 sna, 3 UCINET).
double centralization_max(uint64_t n0, bool directed, int32_t measure, int32_t mode, bool loops,
                          int32_t convention) {
    if (n0 == 0) return NAN;
    const double n = (double)n0;
    if (!directed) mode = 3;
    loops = loops && convention == 0;
    if (measure == 0) {
        if (!directed) return loops ? (n - 1.0) * n : (n - 1.0) * (n - 2.0);
        if (mode == 3) return loops ? 2.0 * (n - 1.0) * (n - 1.0) : 2.0 * (n - 1.0) * (n - 2.0);
        return loops ? (n - 1.0) * n : (n - 1.0) * (n - 1.0);
    }
    if (measure == 1) {
        const double t = (n - 1.0) * (n - 1.0) * (n - 2.0);
        return directed ? t : t / 2.0;
    }
    if (convention == 3 || mode == 3) return (n - 1.0) * (n - 2.0) / (2.0 * n - 3.0);
    return (n - 1.0) * (1.0 - 1.0 / n);
}

// Freeman centralization (topography.rs:376-475): the measure's scores on
// the device, then n * max - sum (in node order; NaN if a score is NaN),
// over the maximum when normalized.
void centralization_entry(const GraphHandle* h, int32_t measure, int32_t mode, int32_t loops,
                          int32_t normalized, int32_t convention, const netbc_options* opt,
                          double* out, netbc_diagnostics* diag) {
    netbc::Graph& g0 = handle(h);
    if (measure < 0 || measure > 3)
        throw netbc::Error("unknown centralization measure code " + std::to_string(measure));
    if (mode < 1 || mode > 3)
        throw netbc::Error("unknown degree mode code " + std::to_string(mode));
    if (convention < 0 || convention > 3)
        throw netbc::Error("unknown convention code " + std::to_string(convention));
    if (convention == 1)
        throw netbc::Error("NetworkX has no centralization index; use the igraph, sna, or "
                           "UCINET convention");
    if (measure == 3) {
        static const char* names[] = {"igraph", "NetworkX", "sna", "UCINET"};
        if (convention != 0)
            throw netbc::Error("eigenvector centralization is provided for the igraph "
                               "convention only, not " + std::string(names[convention]));
        if (g0.directed && mode != 3)
            throw netbc::Error("eigenvector centralization is provided with direction ignored "
                               "(mode \"all\") only");
        throw netbc::Error("eigenvector centralization is not part of this library");
    }
    check_options(opt);
    netbc::Graph simple;
    netbc::Graph* pg = &g0;
    if (convention == 2 || convention == 3) {
        std::lock_guard<std::mutex> lock(g0.mu);
        simplify_graph(g0, false, simple);
        pg = &simple;
    }
    netbc::Graph& g = *pg;
    const uint64_t n = g.n;
    const bool directed = g.directed;
    const int32_t emode = directed ? mode : 3;
    std::vector<double> scores(std::max<uint64_t>(n, 1));
    netbc::RunDiag d;
    double build_ms = 0.0;
    if (measure == 0) {
        if (convention == 3 && directed && emode == 3)
            throw netbc::Error("UCINET reports degree centralization for in- and out-degree only");
        std::lock_guard<std::mutex> lock(g.mu);
        build_ms = prepare(g);
        netbc::DevBuf<double> deg;
        deg.alloc(std::max<uint64_t>(n, 1));
        netbc::degree_scores(g, emode, loops && convention == 0, deg.p);
        NETBC_HIP(hipMemcpyAsync(scores.data(), deg.p, n * sizeof(double), hipMemcpyDeviceToHost,
                                 g.stream));
        NETBC_HIP(hipStreamSynchronize(g.stream));
    } else if (measure == 1) {
        d = run_graph(g, nullptr, (int64_t)n, 0, 0, 0, opt, scores.data(), &build_ms);
    } else {
        netbc::ClosenessRun cr;
        cr.normalized = true;
        cr.unreachable = convention == 2 ? 2 : 0;
        build_ms = closeness_run(g, emode, 0, 0, cr, opt, scores.data(), d);
        // UCINET: NaN unless every node reaches every other
        if (convention == 3 && d.reached_nodes != (int64_t)(n * (n > 0 ? n - 1 : 0))) {
            *out = NAN;
            fill_diag(diag, d, build_ms);
            return;
        }
    }
    double result = NAN;
    if (n > 0) {
        double mx = -INFINITY, total = 0.0;
        bool any_nan = false;
        for (uint64_t v = 0; v < n; v++) {
            const double x = scores[v];
            if (std::isnan(x)) any_nan = true;
            else if (x > mx) mx = x;
            total += x;
        }
        if (any_nan) mx = NAN;
        result = (double)n * mx - total;
        if (normalized)
            result /= centralization_max(n, directed, measure, emode, loops != 0, convention);
    }
    *out = result;
    fill_diag(diag, d, build_ms);
}

// betweenness_max, bipartite.rs:228-241 (Borgatti and Everett 1997): the
// largest betweenness of a node whose mode has `own` nodes, the other mode
// `other`; NaN when either is 0.
double twomode_betweenness_max(uint64_t own, uint64_t other) {
    if (other == 0 || own == 0) return NAN;
    const __int128 m = other, s = (own - 1) / other, t = (own - 1) % other;
    const __int128 twice =
        m * m * (s + 1) * (s + 1) + m * (s + 1) * (2 * t - s - 1) - t * (2 * s - t + 3);
    return (double)twice / 2.0;
}

// Centrality normalized for two-mode data (crates/capi/src/twomode.rs:
// 237-286 -> bipartite.rs:130-269): the checks in the reference's order,
// then the measure on the affiliation graph (undirected, loops dropped,
// parallel edges collapsed) on the device.
void twomode_entry(const GraphHandle* h, const int32_t* types, int32_t measure, int32_t normalized,
                   int32_t unreachable, const netbc_options* opt, double* out,
                   netbc_diagnostics* diag) {
    netbc::Graph& g = handle(h);
    if (!types) throw netbc::Error("node types are missing");
    const uint64_t n = g.n;
    for (uint64_t v = 0; v < n; v++)
        if (types[v] != 0 && types[v] != 1)
            throw netbc::Error("node " + std::to_string(v) + " has type code " +
                               std::to_string(types[v]) +
                               "; use 0 for the first mode and 1 for the second");
    if (measure == 1) {
        if (unreachable != 0 && unreachable != 1)
            throw netbc::Error("unknown code " + std::to_string(unreachable) +
                               " for unreachable nodes");
    } else if (measure != 0 && measure != 2) {
        throw netbc::Error("unknown two-mode centrality code " + std::to_string(measure));
    }
    for (uint64_t e = 0; e < g.m; e++)
        if (types[g.src[e]] == types[g.tgt[e]])
            throw netbc::Error("edge " + std::to_string(e) + " joins nodes " +
                               std::to_string(g.src[e]) + " and " + std::to_string(g.tgt[e]) +
                               ", which are in the same mode; in a two-mode network every edge "
                               "joins the two modes");
    uint64_t second = 0;
    for (uint64_t v = 0; v < n; v++) second += types[v] == 1;
    const uint64_t first = n - second;
    if (measure == 1 && unreachable == 1 && !normalized)
        throw netbc::Error("the Wasserman-Faust rule applies to normalized closeness only");
    check_options(opt);
    netbc::Graph a;
    {
        std::lock_guard<std::mutex> lock(g.mu);
        simplify_graph(g, true, a);
    }
    netbc::RunDiag d;
    double build_ms = prepare(a);
    if (n == 0) {
        fill_diag(diag, d, build_ms);
        return;
    }
    std::vector<uint8_t> tf(n);
    for (uint64_t v = 0; v < n; v++) tf[v] = (uint8_t)types[v];
    netbc::DevBuf<uint8_t> dt;
    dt.alloc(n);
    NETBC_HIP(hipMemcpyAsync(dt.p, tf.data(), n, hipMemcpyHostToDevice, a.stream));
    if (measure == 1) {
        netbc::ClosenessRun cr;
        cr.normalized = normalized != 0;
        cr.unreachable = unreachable;
        cr.d_types = dt.p;
        cr.first = first;
        cr.second = second;
        build_ms += closeness_run(a, 3, 0, 0, cr, opt, out, d);
        fill_diag(diag, d, build_ms);
        return;
    }
    netbc::DevBuf<double> vals;
    vals.alloc(n);
    if (measure == 0) {
        netbc::degree_scores(a, 3, false, vals.p);
        if (normalized) netbc::divide_by_type(a, vals.p, dt.p, (double)second, (double)first);
    } else {
        netbc::Lengths L;
        netbc::classify_lengths(a, false, 0, L);
        netbc::run_batched(a, L, nullptr, (int64_t)n, options(opt), vals.p, d);
        netbc::scale_flagged(a, vals.p, n, nullptr, 0.5, 0.5);  // undirected pairs once
        if (normalized)
            netbc::divide_by_type(a, vals.p, dt.p, twomode_betweenness_max(first, second),
                                  twomode_betweenness_max(second, first));
    }
    NETBC_HIP(hipMemcpyAsync(out, vals.p, n * sizeof(double), hipMemcpyDeviceToHost, a.stream));
    NETBC_HIP(hipStreamSynchronize(a.stream));
    fill_diag(diag, d, build_ms);
}

// Closeness or harmonic centrality (topography.rs:167-199 ->
// closeness.rs:131-209): the option checks in the reference's order, the
// graph oriented for the mode (closeness.rs:225-238: in-paths are the
// out-paths of the reversed graph, paths ignoring direction those of the
// undirected graph), the path lengths, then the profiles and values on the
// device.
void closeness_entry(const GraphHandle* h, int32_t harmonic, int32_t mode, int32_t weighted,
                     int32_t weights_as, int32_t normalized, int32_t unreachable,
                     const netbc_options* opt, double* out, netbc_diagnostics* diag) {
    netbc::Graph& g = handle(h);
    if (mode < 1 || mode > 3)
        throw netbc::Error("unknown degree mode code " + std::to_string(mode));
    check_kind(weights_as);
    if (unreachable < 0 || unreachable > 3)
        throw netbc::Error("unknown code " + std::to_string(unreachable) +
                           " for unreachable nodes");
    if (!harmonic && unreachable == 1 && !normalized)
        throw netbc::Error(
            "the Wasserman-Faust closeness index is defined only in normalized form");
    if (!harmonic && unreachable == 3 && weighted)
        throw netbc::Error(
            "closeness with unreachable lengths set to the longest length plus 1 is defined "
            "only for unweighted networks");
    check_options(opt);
    netbc::ClosenessRun cr;
    cr.harmonic = harmonic != 0;
    cr.normalized = normalized != 0;
    cr.unreachable = unreachable;
    netbc::RunDiag d;
    const double build_ms = closeness_run(g, mode, weighted, weights_as, cr, opt, out, d);
    fill_diag(diag, d, build_ms);
}

// Fills a host graph from edges, with net_graph_new's checks and texts
// (graph.rs:142-234); `prefix` names the graph within a batch.
void fill_graph(netbc::Graph& g, int64_t n, const int64_t* src, const int64_t* tgt,
                const double* weights, int64_t m, bool directed, const std::string& prefix) {
    if (n < 0 || m < 0) throw netbc::Error(prefix + "negative node or edge count");
    if ((uint64_t)n > 0xFFFFFFFFull)
        throw netbc::Error(prefix + "node count (" + std::to_string(n) +
                           ") exceeds the supported maximum");
    if (weights) {
        for (int64_t e = 0; e < m; e++)
            if (!std::isfinite(weights[e]))
                throw netbc::Error(prefix + "edge " + std::to_string(e) +
                                   " has a missing, NaN, or infinite weight");
    }
    if ((uint64_t)m > 0xFFFFFFFFull / 2)
        throw netbc::Error(prefix + "edge count (" + std::to_string(m) +
                           ") exceeds the supported maximum");
    g.n = (uint64_t)n;
    g.m = (uint64_t)m;
    g.directed = directed;
    g.src.resize((size_t)m);
    g.tgt.resize((size_t)m);
    for (int64_t e = 0; e < m; e++) {
        int64_t ends[2] = {src[e], tgt[e]};
        for (int64_t v : ends)
            if (v < 0 || (uint64_t)v >= (uint64_t)n)
                throw netbc::Error(prefix + "edge " + std::to_string(e) +
                                   " refers to node index " + std::to_string(v) +
                                   ", but valid indices are 0.." + std::to_string(n));
        g.src[e] = (uint32_t)src[e];
        g.tgt[e] = (uint32_t)tgt[e];
    }
    if (weights) {
        g.has_weights = true;
        g.w.assign(weights, weights + m);
    }
}

}  // namespace

NETBC_EXPORT int32_t net_graph_new(int64_t n, const int64_t* src, const int64_t* tgt,
                                   const double* weights, int64_t m, int32_t directed,
                                   GraphHandle** out, char* err, size_t err_len) {
    return guard(err, err_len, [&] {
        auto* h = new GraphHandle();
        try {
            fill_graph(h->g, n, src, tgt, weights, m, directed != 0, "");
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

NETBC_EXPORT void net_graph_free(GraphHandle* h) { delete h; }

NETBC_EXPORT int64_t net_graph_node_count(const GraphHandle* h) {
    return h ? (int64_t)h->g.n : -1;
}

NETBC_EXPORT int64_t net_graph_edge_count(const GraphHandle* h) {
    return h ? (int64_t)h->g.m : -1;
}

NETBC_EXPORT int32_t net_betweenness(const GraphHandle* h, int32_t weighted, int32_t weights_as,
                                     int32_t normalized, int32_t threads, double* out, char* err,
                                     size_t err_len) {
    (void)threads;
    return guard(err, err_len, [&] {
        netbc::Graph& g = handle(h);
        check_kind(weights_as);
        run_sources(g, nullptr, (int64_t)g.n, weighted, weights_as, normalized, nullptr, out,
                    nullptr);
    });
}

NETBC_EXPORT int32_t netbc_betweenness_sources(const GraphHandle* h, const int64_t* sources,
                                               int64_t k, int32_t weighted, int32_t weights_as,
                                               int32_t normalized, const netbc_options* opt,
                                               double* out, netbc_diagnostics* diag, char* err,
                                               size_t err_len) {
    return guard(err, err_len, [&] {
        netbc::Graph& g = handle(h);
        check_kind(weights_as);
        k = check_sources(g, sources, k);
        check_options(opt);
        run_sources(g, sources, k, weighted, weights_as, normalized, opt, out, diag);
    });
}

NETBC_EXPORT int32_t net_edge_betweenness(const GraphHandle* h, int32_t weighted,
                                          int32_t weights_as, int32_t normalized, int32_t threads,
                                          double* out, char* err, size_t err_len) {
    (void)threads;
    return guard(err, err_len, [&] {
        netbc::Graph& g = handle(h);
        check_kind(weights_as);
        run_sources(g, nullptr, (int64_t)g.n, weighted, weights_as, normalized, nullptr, out,
                    nullptr, true);
    });
}

NETBC_EXPORT int32_t net_betweenness_subset(const GraphHandle* h, const int64_t* sources,
                                            int64_t n_sources, const int64_t* targets,
                                            int64_t n_targets, int32_t weighted,
                                            int32_t weights_as, int32_t endpoints,
                                            int32_t normalized, int32_t rescale,
                                            int32_t parallel, int32_t edge_split, int32_t tie,
                                            int32_t threads, double* nodes, double* edges,
                                            char* err, size_t err_len) {
    (void)threads;
    return guard(err, err_len, [&] {
        subset_entry(h, sources, n_sources, targets, n_targets, weighted, weights_as, endpoints,
                     normalized, rescale, parallel, edge_split, tie, nullptr, nodes, edges,
                     nullptr);
    });
}

NETBC_EXPORT int32_t netbc_betweenness_subset_ex(
    const GraphHandle* h, const int64_t* sources, int64_t n_sources, const int64_t* targets,
    int64_t n_targets, int32_t weighted, int32_t weights_as, int32_t endpoints,
    int32_t normalized, int32_t rescale, int32_t parallel, int32_t edge_split, int32_t tie,
    const netbc_options* opt, double* nodes, double* edges, netbc_diagnostics* diag, char* err,
    size_t err_len) {
    return guard(err, err_len, [&] {
        subset_entry(h, sources, n_sources, targets, n_targets, weighted, weights_as, endpoints,
                     normalized, rescale, parallel, edge_split, tie, opt, nodes, edges, diag);
    });
}

NETBC_EXPORT int32_t net_closeness(const GraphHandle* h, int32_t harmonic, int32_t mode,
                                   int32_t weighted, int32_t weights_as, int32_t normalized,
                                   int32_t unreachable, int32_t threads, double* out, char* err,
                                   size_t err_len) {
    (void)threads;
    return guard(err, err_len, [&] {
        closeness_entry(h, harmonic, mode, weighted, weights_as, normalized, unreachable, nullptr,
                        out, nullptr);
    });
}

NETBC_EXPORT int32_t netbc_closeness_ex(const GraphHandle* h, int32_t harmonic, int32_t mode,
                                        int32_t weighted, int32_t weights_as, int32_t normalized,
                                        int32_t unreachable, const netbc_options* opt,
                                        double* out, netbc_diagnostics* diag, char* err,
                                        size_t err_len) {
    return guard(err, err_len, [&] {
        closeness_entry(h, harmonic, mode, weighted, weights_as, normalized, unreachable, opt,
                        out, diag);
    });
}

NETBC_EXPORT int32_t net_centralization(const GraphHandle* h, int32_t measure, int32_t mode,
                                        int32_t loops, int32_t normalized, int32_t convention,
                                        int32_t threads, double* out, char* err,
                                        size_t err_len) {
    (void)threads;
    return guard(err, err_len, [&] {
        centralization_entry(h, measure, mode, loops, normalized, convention, nullptr, out,
                             nullptr);
    });
}

NETBC_EXPORT int32_t net_twomode_centrality(const GraphHandle* h, const int32_t* types,
                                            int32_t measure, int32_t normalized,
                                            int32_t unreachable, int32_t threads, double* out,
                                            char* err, size_t err_len) {
    (void)threads;
    return guard(err, err_len, [&] {
        twomode_entry(h, types, measure, normalized, unreachable, nullptr, out, nullptr);
    });
}

NETBC_EXPORT int32_t netbc_edge_betweenness_sources(const GraphHandle* h, const int64_t* sources,
                                                    int64_t k, int32_t weighted,
                                                    int32_t weights_as, int32_t normalized,
                                                    const netbc_options* opt, double* out,
                                                    netbc_diagnostics* diag, char* err,
                                                    size_t err_len) {
    return guard(err, err_len, [&] {
        netbc::Graph& g = handle(h);
        check_kind(weights_as);
        k = check_sources(g, sources, k);
        check_options(opt);
        run_sources(g, sources, k, weighted, weights_as, normalized, opt, out, diag, true);
    });
}

NETBC_EXPORT int32_t netbc_state_dump(const GraphHandle* h, int32_t weighted, int32_t weights_as,
                                      int32_t precision, const int64_t* sources, int64_t k,
                                      uint64_t* dist, double* sigma, double* delta, char* err,
                                      size_t err_len) {
    return guard(err, err_len, [&] {
        netbc::Graph& g = handle(h);
        check_kind(weights_as);
        for (int64_t i = 0; i < k; i++)
            if (sources[i] < 0 || (uint64_t)sources[i] >= g.n)
                throw netbc::Error("source out of range");
        std::lock_guard<std::mutex> lock(g.mu);
        if (weighted && !g.has_weights) throw netbc::Error("the graph has no edge weights");
        prepare(g);
        netbc::Lengths L;
        netbc::classify_lengths(g, weighted != 0, weights_as, L);
        netbc::RunOptions o;
        o.precision = precision;
        netbc::run_dump(g, L, sources, k, o, dist, sigma, delta);
    });
}

NETBC_EXPORT int64_t netbc_graph_arc_count(const GraphHandle* h) {
    if (!h) return -1;
    const netbc::Graph& g = h->g;
    if (g.directed) return (int64_t)g.m;
    int64_t loops = 0;
    for (uint64_t e = 0; e < g.m; e++) loops += g.src[e] == g.tgt[e];
    return (int64_t)(2 * g.m) - loops;
}

NETBC_EXPORT int32_t netbc_graph_csr(const GraphHandle* h, int32_t incoming, uint32_t* offsets,
                                     uint32_t* targets, double* weights, char* err,
                                     size_t err_len) {
    return guard(err, err_len, [&] {
        netbc::Graph& g = handle(h);
        std::lock_guard<std::mutex> lock(g.mu);
        prepare(g);
        const netbc::DeviceGraph& d = g.dev;
        bool in = incoming && d.directed;
        const uint32_t* off = in ? d.in_off.p : d.out_off.p;
        const uint32_t* idx = in ? d.in_idx.p : d.out_idx.p;
        const double* w = in ? d.in_w.p : d.out_w.p;
        NETBC_HIP(hipMemcpy(offsets, off, (g.n + 1) * 4, hipMemcpyDeviceToHost));
        if (d.arcs) NETBC_HIP(hipMemcpy(targets, idx, d.arcs * 4, hipMemcpyDeviceToHost));
        if (weights && d.weighted && d.arcs)
            NETBC_HIP(hipMemcpy(weights, w, d.arcs * 8, hipMemcpyDeviceToHost));
    });
}

NETBC_EXPORT int32_t netbc_device_info(char* buf, size_t len) {
    return guard(buf, len, [&] {
        ensure_device();
        int dev = 0;
        NETBC_HIP(hipGetDevice(&dev));
        hipDeviceProp_t p;
        NETBC_HIP(hipGetDeviceProperties(&p, dev));
        std::string s = std::string(p.name) + ", " + p.gcnArchName + ", " +
                        std::to_string(p.multiProcessorCount) + " CUs";
        write_err(buf, len, s);
    });
}

namespace {

// The one-workgroup engine's layout of a batch (small.h): a descriptor per
// graph (source-group size and count as parallel.rs:437-440, scale as
// betweenness.rs:100-113, group sums one graph after another) and the
// graph of every union node. Returns whether the engine takes any graph
// (unweighted, 1..2048 nodes).
bool small_layout(int64_t graphs, const int64_t* node_counts, const std::vector<uint64_t>& node0,
                  bool directed, bool normalized, bool weighted,
                  std::vector<netbc::SmallGraph>& descs, std::vector<uint32_t>& node_graph) {
    descs.assign((size_t)graphs, netbc::SmallGraph{});
    node_graph.assign(node0[graphs], 0);
    uint64_t gsum_off = 0;
    bool any_device = false;
    for (int64_t i = 0; i < graphs; i++) {
        netbc::SmallGraph& G = descs[i];
        uint64_t n = (uint64_t)node_counts[i];
        G.n = (uint32_t)n;
        G.node0 = (uint32_t)node0[i];
        G.gs = n ? (uint32_t)((n + 63) / 64) : 1;
        G.groups = n ? (uint32_t)((n + G.gs - 1) / G.gs) : 0;
        G.scale = graph_scale(n, directed, normalized);
        G.device = (!weighted && n > 0 && n <= netbc::kSmallMaxNodes) ? 1 : 0;
        G.gsum_off = gsum_off;
        G.pad = 0;
        if (G.device) {
            gsum_off += (uint64_t)G.groups * n;
            any_device = true;
        } else {
            G.groups = 0;
        }
        for (uint64_t v = 0; v < n; v++) node_graph[node0[i] + v] = (uint32_t)i;
    }
    return any_device;
}

// Graphs the one-workgroup engine did not take (or whose search went too
// deep): the batched engine, one graph at a time.
void run_left_over(int64_t graphs, const int64_t* node_counts, const int64_t* edge_counts,
                   const std::vector<uint64_t>& node0, const std::vector<uint64_t>& edge0,
                   const int64_t* src, const int64_t* tgt, const double* weights,
                   int32_t directed, int32_t weighted, int32_t weights_as, int32_t normalized,
                   const netbc_options* opt, const std::vector<netbc::SmallGraph>& descs,
                   const std::vector<uint32_t>& too_deep, double* out, netbc::RunDiag& total) {
    for (int64_t i = 0; i < graphs; i++) {
        if (descs[i].device && !too_deep[i]) continue;
        if (node_counts[i] == 0) continue;
        netbc::Graph one;
        uint64_t e0 = edge0[i];
        fill_graph(one, node_counts[i], src + e0, tgt + e0, weights ? weights + e0 : nullptr,
                   edge_counts[i], directed != 0, "");
        netbc::RunDiag d;
        try {
            d = run_graph(one, nullptr, node_counts[i], weighted, weights_as, normalized, opt,
                          out + node0[i], nullptr);
        } catch (const netbc::Error& e) {
            throw netbc::Error("graph " + std::to_string(i) + ": " + e.what());
        }
        merge_diag(total, d);
    }
}

}  // namespace

NETBC_EXPORT int32_t netbc_betweenness_batch(int64_t graphs, const int64_t* node_counts,
                                             const int64_t* edge_counts, const int64_t* src,
                                             const int64_t* tgt, const double* weights,
                                             int32_t directed, int32_t weighted,
                                             int32_t weights_as, int32_t normalized,
                                             const netbc_options* opt, double* out,
                                             netbc_diagnostics* diag, char* err,
                                             size_t err_len) {
    return guard(err, err_len, [&] {
        if (graphs < 0) throw netbc::Error("negative graph count");
        check_kind(weights_as);
        check_options(opt);
        if (weighted && !weights) throw netbc::Error("the graph has no edge weights");
        // Validate every graph as net_graph_new would, and lay them out as
        // one union graph with consecutive node ranges.
        std::vector<uint64_t> node0((size_t)graphs + 1, 0), edge0((size_t)graphs + 1, 0);
        for (int64_t i = 0; i < graphs; i++) {
            if (node_counts[i] < 0 || edge_counts[i] < 0)
                throw netbc::Error("graph " + std::to_string(i) +
                                   ": negative node or edge count");
            node0[i + 1] = node0[i] + (uint64_t)node_counts[i];
            edge0[i + 1] = edge0[i] + (uint64_t)edge_counts[i];
        }
        if (node0[graphs] > 0xFFFFFFFFull || edge0[graphs] > 0xFFFFFFFFull / 2)
            throw netbc::Error("the batch exceeds 2^32 nodes or 2^31 edges in total");
        netbc::Graph U;
        U.n = node0[graphs];
        U.m = edge0[graphs];
        U.directed = directed != 0;
        U.src.resize(U.m);
        U.tgt.resize(U.m);
        if (weights) {
            U.has_weights = true;
            U.w.assign(weights, weights + U.m);
        }
        // one pass over the edges: each graph's checks (fill_graph's, with
        // its messages) and its arcs shifted to its node range
        for (int64_t i = 0; i < graphs; i++) {
            const uint64_t e0 = edge0[i], m = (uint64_t)edge_counts[i];
            const uint64_t n = (uint64_t)node_counts[i];
            const uint32_t shift = (uint32_t)node0[i];
            auto fail = [&](const std::string& what) {
                throw netbc::Error("graph " + std::to_string(i) + ": " + what);
            };
            if (n > 0xFFFFFFFFull)
                fail("node count (" + std::to_string(n) + ") exceeds the supported maximum");
            if (weights)
                for (uint64_t e = 0; e < m; e++)
                    if (!std::isfinite(weights[e0 + e]))
                        fail("edge " + std::to_string(e) +
                             " has a missing, NaN, or infinite weight");
            if (m > 0xFFFFFFFFull / 2)
                fail("edge count (" + std::to_string(m) + ") exceeds the supported maximum");
            const int64_t* si = src + e0;
            const int64_t* ti = tgt + e0;
            uint32_t* su = U.src.data() + e0;
            uint32_t* tu = U.tgt.data() + e0;
            for (uint64_t e = 0; e < m; e++) {
                const int64_t a = si[e], b = ti[e];
                if ((uint64_t)a >= n || (uint64_t)b >= n) {
                    const int64_t v = (uint64_t)a >= n ? a : b;
                    fail("edge " + std::to_string(e) + " refers to node index " +
                         std::to_string(v) + ", but valid indices are 0.." + std::to_string(n));
                }
                su[e] = (uint32_t)a + shift;
                tu[e] = (uint32_t)b + shift;
            }
        }
        netbc::RunDiag total;
        netbc::RunOptions o = options(opt);
        std::vector<netbc::SmallGraph> descs;
        std::vector<uint32_t> node_graph;
        const bool any_device = small_layout(graphs, node_counts, node0, directed != 0,
                                             normalized != 0, weighted != 0, descs, node_graph);
        std::vector<uint32_t> too_deep((size_t)graphs, 0);
        if (any_device) {
            ensure_device();
            netbc::build_device_csr(U);
            netbc::run_small_batch(U, descs, node_graph, o, out, too_deep, total);
        }
        run_left_over(graphs, node_counts, edge_counts, node0, edge0, src, tgt, weights,
                      directed, weighted, weights_as, normalized, opt, descs, too_deep, out,
                      total);
        fill_diag(diag, total, 0.0);
    });
}

NETBC_EXPORT int32_t netbc_betweenness_batch_noalloc(
    int64_t graphs, const int64_t* node_counts, int32_t directed, int32_t normalized,
    const uint32_t* offsets, const uint32_t* targets, const uint32_t* in_offsets,
    const uint32_t* in_targets, int64_t arcs, double* out, void* stream, char* err,
    size_t err_len) {
    return guard(err, err_len, [&] {
        ensure_device();
        netbc::RunDiag d;
        netbc::run_small_batch_noalloc(graphs, node_counts, directed != 0, normalized != 0,
                                       offsets, targets, in_offsets, in_targets, arcs, out,
                                       static_cast<hipStream_t>(stream), d);
    });
}

// ---- batch plans: one batch shape, replayed ------------------------------

struct netbc_batch_plan {
    int device = 0;
    hipStream_t st = nullptr;
    bool directed = false, normalized = false, captured = false, any_device = false;
    int64_t graphs = 0;
    bool has_opt = false;
    netbc_options opt{};
    std::vector<int64_t> node_counts, edge_counts;
    std::vector<uint64_t> node0, edge0;
    std::vector<netbc::SmallGraph> descs;
    netbc::FixedCsr csr;
    netbc::SmallRun run;
    uint32_t* h_src = nullptr;  // pinned staging of the union's edges
    uint32_t* h_tgt = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::mutex mu;

    ~netbc_batch_plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (h_src) (void)hipHostFree(h_src);
        if (h_tgt) (void)hipHostFree(h_tgt);
        if (st) (void)hipStreamDestroy(st);
    }
    const netbc_options* options_p() const { return has_opt ? &opt : nullptr; }
    // the device side of one run: the union CSR from the staged edges, then
    // the persistent launch and the reduction
    void enqueue() {
        netbc::fixed_csr_enqueue(csr, st);
        netbc::small_enqueue(run, csr.out_off.p, csr.out_idx.p, csr.in_off_p(), csr.in_idx_p(),
                             st);
    }
};

namespace {
// Makes `dev` current for the scope and restores the caller's device.
struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev) {
        NETBC_HIP(hipGetDevice(&prev));
        if (prev != dev) NETBC_HIP(hipSetDevice(dev));
    }
    ~DeviceScope() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};
}  // namespace

NETBC_EXPORT int32_t netbc_batch_plan_new(int64_t graphs, const int64_t* node_counts,
                                          const int64_t* edge_counts, int32_t directed,
                                          int32_t normalized, int32_t capture,
                                          const netbc_options* opt, netbc_batch_plan** plan,
                                          char* err, size_t err_len) {
    return guard(err, err_len, [&] {
        if (!plan) throw netbc::Error("null plan pointer");
        *plan = nullptr;
        ensure_device();
        if (graphs < 0) throw netbc::Error("negative graph count");
        if (graphs > 0 && (!node_counts || !edge_counts))
            throw netbc::Error("null node or edge counts");
        check_options(opt);
        auto P = std::make_unique<netbc_batch_plan>();
        NETBC_HIP(hipGetDevice(&P->device));
        P->graphs = graphs;
        P->directed = directed != 0;
        P->normalized = normalized != 0;
        if (opt) {
            P->has_opt = true;
            P->opt = *opt;
        }
        P->node_counts.assign(node_counts, node_counts + graphs);
        P->edge_counts.assign(edge_counts, edge_counts + graphs);
        P->node0.assign((size_t)graphs + 1, 0);
        P->edge0.assign((size_t)graphs + 1, 0);
        for (int64_t i = 0; i < graphs; i++) {
            const std::string prefix = "graph " + std::to_string(i) + ": ";
            if (node_counts[i] < 0 || edge_counts[i] < 0)
                throw netbc::Error(prefix + "negative node or edge count");
            if ((uint64_t)node_counts[i] > 0xFFFFFFFFull)
                throw netbc::Error(prefix + "node count (" + std::to_string(node_counts[i]) +
                                   ") exceeds the supported maximum");
            if ((uint64_t)edge_counts[i] > 0xFFFFFFFFull / 2)
                throw netbc::Error(prefix + "edge count (" + std::to_string(edge_counts[i]) +
                                   ") exceeds the supported maximum");
            P->node0[i + 1] = P->node0[i] + (uint64_t)node_counts[i];
            P->edge0[i + 1] = P->edge0[i] + (uint64_t)edge_counts[i];
        }
        const uint64_t n = P->node0[graphs], m = P->edge0[graphs];
        if (n > 0xFFFFFFFFull || m > 0xFFFFFFFFull / 2)
            throw netbc::Error("the batch exceeds 2^32 nodes or 2^31 edges in total");
        std::vector<uint32_t> node_graph;
        P->any_device = small_layout(graphs, node_counts, P->node0, P->directed, P->normalized,
                                     false, P->descs, node_graph);
        if (!P->any_device) {
            *plan = P.release();
            return;
        }
        NETBC_HIP(hipStreamCreateWithFlags(&P->st, hipStreamNonBlocking));
        NETBC_HIP(hipEventCreate(&P->e0));
        NETBC_HIP(hipEventCreate(&P->e1));
        NETBC_HIP(hipHostMalloc((void**)&P->h_src, std::max<uint64_t>(m, 1) * 4));
        NETBC_HIP(hipHostMalloc((void**)&P->h_tgt, std::max<uint64_t>(m, 1) * 4));
        netbc::fixed_csr_alloc(P->csr, n, m, P->directed);
        netbc::small_setup(P->run, P->descs, node_graph, options(opt), P->st);
        if (capture) {
            NETBC_HIP(hipStreamBeginCapture(P->st, hipStreamCaptureModeThreadLocal));
            try {
                P->enqueue();
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(P->st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            NETBC_HIP(hipStreamEndCapture(P->st, &P->graph));
            NETBC_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
            P->captured = true;
        }
        *plan = P.release();
    });
}

NETBC_EXPORT int32_t netbc_batch_plan_run(netbc_batch_plan* plan, const int64_t* src,
                                          const int64_t* tgt, double* out,
                                          netbc_diagnostics* diag, char* err, size_t err_len) {
    return guard(err, err_len, [&] {
        if (!plan) throw netbc::Error("plan is null");
        netbc_batch_plan& P = *plan;
        std::lock_guard<std::mutex> lock(P.mu);
        ensure_device();
        const uint64_t m = P.edge0[P.graphs];
        if (m && (!src || !tgt)) throw netbc::Error("null edge arrays");
        if (P.node0[P.graphs] && !out) throw netbc::Error("null score array");
        netbc::RunDiag total;
        std::vector<uint32_t> too_deep((size_t)P.graphs, 0);
        // validate as net_graph_new does and stage as union node ids
        for (int64_t i = 0; i < P.graphs; i++) {
            const int64_t n = P.node_counts[i];
            const uint64_t e0 = P.edge0[i];
            const uint32_t base = (uint32_t)P.node0[i];
            for (int64_t e = 0; e < P.edge_counts[i]; e++) {
                const int64_t s = src[e0 + e], t = tgt[e0 + e];
                const int64_t ends[2] = {s, t};
                for (int64_t v : ends)
                    if (v < 0 || v >= n)
                        throw netbc::Error("graph " + std::to_string(i) + ": edge " +
                                           std::to_string(e) + " refers to node index " +
                                           std::to_string(v) + ", but valid indices are 0.." +
                                           std::to_string(n));
                if (P.h_src) {
                    P.h_src[e0 + e] = (uint32_t)s + base;
                    P.h_tgt[e0 + e] = (uint32_t)t + base;
                }
            }
        }
        if (P.any_device) {
            DeviceScope scope(P.device);
            if (m) {
                NETBC_HIP(hipMemcpyAsync(P.csr.src.p, P.h_src, m * 4, hipMemcpyHostToDevice,
                                         P.st));
                NETBC_HIP(hipMemcpyAsync(P.csr.tgt.p, P.h_tgt, m * 4, hipMemcpyHostToDevice,
                                         P.st));
            }
            NETBC_HIP(hipEventRecord(P.e0, P.st));
            if (P.captured) NETBC_HIP(hipGraphLaunch(P.exec, P.st));
            else P.enqueue();
            NETBC_HIP(hipEventRecord(P.e1, P.st));
            unsigned long long hs[2] = {0, 0};
            netbc::small_fetch(P.run, out, too_deep, hs, P.st);
            float ms = 0.0f;
            NETBC_HIP(hipEventElapsedTime(&ms, P.e0, P.e1));
            total.kernel_ms += ms;
            total.reached_nodes += (int64_t)hs[0];
            total.reached_arcs += (int64_t)hs[1];
            total.slots = (int32_t)P.run.grid;
        }
        run_left_over(P.graphs, P.node_counts.data(), P.edge_counts.data(), P.node0, P.edge0,
                      src, tgt, nullptr, P.directed ? 1 : 0, 0, 0, P.normalized ? 1 : 0,
                      P.options_p(), P.descs, too_deep, out, total);
        fill_diag(diag, total, 0.0);
    });
}

NETBC_EXPORT void netbc_batch_plan_free(netbc_batch_plan* plan) { delete plan; }

namespace {
// A plain streaming copy (16 bytes per thread per iteration).
__global__ void netbc_copy_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                  size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) dst[i] = src[i];
}

// Random 8-byte reads: one word at the start of each of `lines` lines of
// `line` bytes, every line once, in the scattered order i * 0x9E3779B1
// mod lines (a bijection: lines is a power of two and the factor odd).
// Each thread writes its sum to dst[thread], so the kernel's writes are
// grid-sized and its reads are the gathers alone.
__global__ void netbc_gather_kernel(unsigned long long* __restrict__ dst,
                                    const unsigned char* __restrict__ src, size_t lines,
                                    size_t line) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    unsigned long long s = 0;
    for (size_t i = tid; i < lines; i += stride) {
        const size_t l = (i * 0x9E3779B1ull) & (lines - 1);
        s += *reinterpret_cast<const unsigned long long*>(src + l * line);
    }
    dst[tid] = s;
}
}  // namespace

NETBC_EXPORT int32_t netbc_calibration_copy(void* dst, const void* src, size_t bytes, char* err,
                                            size_t err_len) {
    return guard(err, err_len, [&] {
        ensure_device();
        if (bytes % 16) throw netbc::Error("bytes must be a multiple of 16");
        netbc::check_device_range(dst, bytes, "destination");
        netbc::check_device_range(src, bytes, "source");
        size_t count = bytes / 16;
        netbc_copy_kernel<<<2048, 256>>>((uint4*)dst, (const uint4*)src, count);
        NETBC_HIP(hipGetLastError());
        NETBC_HIP(hipDeviceSynchronize());
    });
}

NETBC_EXPORT int32_t netbc_calibration_gather(void* dst, const void* src, size_t bytes,
                                              size_t line, char* err, size_t err_len) {
    return guard(err, err_len, [&] {
        ensure_device();
        if (line < 8 || (line & (line - 1)))
            throw netbc::Error("line must be a power of two of at least 8 bytes");
        if (bytes < line || bytes % line) throw netbc::Error("bytes must be a multiple of line");
        const size_t lines = bytes / line;
        if (lines & (lines - 1)) throw netbc::Error("bytes / line must be a power of two");
        constexpr unsigned kGrid = 2048, kThreads = 256;
        netbc::check_device_range(dst, (size_t)kGrid * kThreads * 8, "destination");
        netbc::check_device_range(src, bytes, "source");
        netbc_gather_kernel<<<kGrid, kThreads>>>((unsigned long long*)dst,
                                                 (const unsigned char*)src, lines, line);
        NETBC_HIP(hipGetLastError());
        NETBC_HIP(hipDeviceSynchronize());
    });
}
