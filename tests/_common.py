"""Helpers shared by the tests (random graphs, tolerance rule)."""
import numpy as np

# validation/compare.py:313-314: |a - b| <= abs + rel * max(|a|, |b|).
def close(a, b, rel=1e-9, abs_=1e-9):
    """NaN matches NaN (a normalized score of a one-node graph is NaN)."""
    a = np.asarray([np.nan if x is None else x for x in np.ravel(a)], dtype=np.float64)
    b = np.asarray([np.nan if x is None else x for x in np.ravel(b)], dtype=np.float64)
    if a.shape != b.shape:
        return False
    nan = np.isnan(a) | np.isnan(b)
    if not np.array_equal(np.isnan(a), np.isnan(b)):
        return False
    a, b = a[~nan], b[~nan]
    return bool(np.all(np.abs(a - b) <= abs_ + rel * np.maximum(np.abs(a), np.abs(b))))


def random_case(rng, max_n=40, loops=True):
    """A random multigraph as crates/engine/tests/properties.rs:30-48 draws
    them: n <= max_n, m up to 3n, loops and parallel edges allowed."""
    n = int(rng.integers(1, max_n + 1))
    m = int(rng.integers(0, 3 * n + 1))
    src = rng.integers(0, n, size=m)
    tgt = rng.integers(0, n, size=m)
    if not loops:
        keep = src != tgt
        src, tgt = src[keep], tgt[keep]
    directed = bool(rng.integers(0, 2))
    return n, src.astype(np.int64), tgt.astype(np.int64), directed


VARIANTS = {
    "betweenness": dict(weighted=False, normalized=False),
    "betweenness_normalized": dict(weighted=False, normalized=True),
    "betweenness_weighted": dict(weighted=True, normalized=False),
    "betweenness_weighted_strength": dict(weighted=True, normalized=False,
                                          weights_as="strength"),
}


EDGE_VARIANTS = {
    "edge_betweenness": dict(weighted=False, normalized=False),
    "edge_betweenness_normalized": dict(weighted=False, normalized=True),
    "edge_betweenness_weighted": dict(weighted=True, normalized=False),
    "edge_betweenness_weighted_strength": dict(weighted=True, normalized=False,
                                               weights_as="strength"),
}


def tolerance(variant):
    # validation/manifest.d/core.yaml:86, 107, 127, 140 and
    # validation/manifest.d/roles.yaml:131, 151, 166, 179
    normalized = variant in ("betweenness_normalized", "edge_betweenness_normalized")
    return (1e-9, 1e-12) if normalized else (1e-9, 1e-9)


# NetworkX counts parallel edges as one path (and credits each parallel edge
# with the whole amount), compares reciprocal lengths exactly and leaves a
# one-node graph unscaled (validation/manifest.d/core.yaml:90-96, 134-140;
# roles.yaml:139-145, 156-159, 181-189).
def networkx_documented(fx, variant):
    tags = set(fx["tags"])
    if "multi_edges" in tags and variant in ("betweenness", "betweenness_normalized",
                                             "edge_betweenness",
                                             "edge_betweenness_normalized"):
        return True
    if variant.endswith("weighted_strength") and "integer_weights" in tags:
        return True
    if variant == "edge_betweenness_normalized" and fx["n"] < 2:
        return True
    return False


# Betweenness for sources and targets (validation/manifest.d/centrality_extra.yaml:
# cx_betweenness_subset; node sets as validation/bindings/centrality_extra.py:
# 60-63, 140-155). rescale 0 unordered, 1 NetworkX, 2 NetworkX subset;
# parallel 0 distinct, 1 collapse; edge_split 0 proportional, 1 NetworkX;
# tie 0 relative, 1 exact.
_NX_SUBSET = dict(sources="even", targets="thirds", rescale=2, parallel=1, tie=1)
_NX_SAMPLE = dict(sample=7, seed=3, rescale=1, parallel=1, tie=1)
SUBSET_VARIANTS = {
    "cx_bc_all_distinct": dict(out="nodes", rescale=0, parallel=0),
    "cx_bc_all_distinct_weighted": dict(out="nodes", rescale=0, parallel=0, weighted=True),
    "cx_ebc_all_distinct": dict(out="edges", rescale=0, parallel=0),
    "cx_bc_subset_nx": dict(out="nodes", **_NX_SUBSET),
    "cx_bc_subset_nx_normalized": dict(out="nodes", normalized=True, **_NX_SUBSET),
    "cx_bc_subset_nx_weighted": dict(out="nodes", weighted=True, **_NX_SUBSET),
    "cx_ebc_subset_nx": dict(out="edges", edge_split=1, **_NX_SUBSET),
    "cx_ebc_subset_nx_weighted": dict(out="edges", edge_split=1, weighted=True, **_NX_SUBSET),
    "cx_ebc_subset_proportional": dict(out="edges", edge_split=0, **_NX_SUBSET),
    "cx_bc_sampled": dict(out="nodes", normalized=True, **_NX_SAMPLE),
    "cx_bc_sampled_raw": dict(out="nodes", **_NX_SAMPLE),
    "cx_bc_sampled_endpoints": dict(out="nodes", normalized=True, endpoints=True, **_NX_SAMPLE),
    "cx_bc_sampled_weighted": dict(out="nodes", normalized=True, weighted=True, **_NX_SAMPLE),
    "cx_ebc_sampled": dict(out="edges", normalized=True, **_NX_SAMPLE),
}


def subset_call(n, spec):
    """(keyword arguments of a betweenness_subset call, which output) for a spec."""
    import random
    kw = {k: v for k, v in spec.items() if k not in ("out", "sample", "seed", "sources",
                                                       "targets")}

    def node_set(kind):
        if kind == "even":
            return [v for v in range(n) if v % 2 == 0]
        if kind == "thirds":
            return [v for v in range(n) if v % 3 == 0]
        return None

    if "sample" in spec:
        kw["sources"] = random.Random(spec["seed"]).sample(list(range(n)), min(spec["sample"], n))
    else:
        kw["sources"] = node_set(spec.get("sources", "all"))
    kw["targets"] = node_set(spec.get("targets", "all"))
    return kw, spec["out"]


def golden_values(vals):
    """The centrality_extra goldens write NaN as -999 and infinity as -1000."""
    out = []
    for x in vals:
        out.append(float("nan") if x is None or x == -999 else (float("inf") if x == -1000 else x))
    return out


def subset_networkx_documented(fx, variant):
    # centrality_extra.yaml:16-58: the all_distinct variants count parallel
    # edges as distinct paths, NetworkX once.
    return "multi_edges" in set(fx["tags"]) and "all_distinct" in variant


# Closeness and harmonic centrality (validation/manifest.d/topography.yaml:
# 151-300; runner validation/bindings/topography.py:130-151). mode 1 out, 2
# in, 3 all; unreachable 0 ignore, 1 Wasserman-Faust, 2 zero, 3 max plus
# one. The `simple: true` variants read the graph as an adjacency matrix,
# which changes no unweighted distance, so they run on the graph itself.
# Tolerance rel 1e-12, abs 1e-14; NaN is written as -1.
CLOSENESS_VARIANTS = {
    "closeness": dict(mode=1),
    "closeness_normalized": dict(mode=1, normalized=True),
    "closeness_in": dict(mode=2, normalized=True),
    "closeness_all": dict(mode=3, normalized=True),
    "closeness_wf": dict(mode=2, normalized=True, unreachable=1),
    "closeness_sna": dict(mode=1, normalized=True, unreachable=2),
    "closeness_ucinet": dict(mode=1, normalized=True, unreachable=3),
    "closeness_ucinet_in": dict(mode=2, normalized=True, unreachable=3),
    "closeness_weighted": dict(mode=1, weighted=True, normalized=True),
    "closeness_wf_weighted": dict(mode=2, weighted=True, normalized=True, unreachable=1),
    "closeness_weighted_strength": dict(mode=1, weighted=True, weights_as="strength",
                                        normalized=True),
    "harmonic": dict(harmonic=True, mode=1),
    "harmonic_normalized": dict(harmonic=True, mode=1, normalized=True),
    "harmonic_in": dict(harmonic=True, mode=2),
    "harmonic_weighted": dict(harmonic=True, mode=1, weighted=True),
}


def closeness_golden(vals):
    return [float("nan") if x is None or x == -1 else x for x in vals]


# Freeman centralization (validation/manifest.d/topography.yaml; the
# validation bindings' defaults: mode all, loops counted, normalized, igraph)
CENTRALIZATION_VARIANTS = {
    "centralization_degree": dict(measure=0, mode=3, loops=True, convention=0),
    "centralization_degree_in": dict(measure=0, mode=2, loops=True, convention=0),
    "centralization_degree_sna": dict(measure=0, mode=3, convention=2),
    "centralization_degree_ucinet": dict(measure=0, mode=1, convention=3),
    "centralization_betweenness": dict(measure=1, convention=0),
    "centralization_closeness": dict(measure=2, mode=1, convention=0),
    "centralization_closeness_sna": dict(measure=2, mode=1, convention=2),
    "centralization_closeness_ucinet": dict(measure=2, mode=1, convention=3),
}

# centrality normalized for two-mode data (validation/manifest.d/twomode.yaml;
# defaults normalized, unreachable nodes ignored)
TWOMODE_VARIANTS = {
    "twomode_degree": dict(measure=0, normalized=True),
    "twomode_degree_raw": dict(measure=0, normalized=False),
    "twomode_closeness": dict(measure=1, normalized=True, unreachable=1),
    "twomode_closeness_ignore": dict(measure=1, normalized=True, unreachable=0),
    "twomode_closeness_raw": dict(measure=1, normalized=False, unreachable=0),
    "twomode_betweenness": dict(measure=2, normalized=True),
    "twomode_betweenness_raw": dict(measure=2, normalized=False),
}
