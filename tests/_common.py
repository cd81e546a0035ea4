"""Helpers shared by the tests (random graphs, tolerance rule)."""
import numpy as np

# validation/compare.py:313-314: |a - b| <= abs + rel * max(|a|, |b|).
def close(a, b, rel=1e-9, abs_=1e-9):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return False
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


def tolerance(variant):
    # validation/manifest.d/core.yaml:86, 107, 127, 140
    return (1e-9, 1e-12) if variant == "betweenness_normalized" else (1e-9, 1e-9)


# NetworkX counts parallel edges as one path and compares reciprocal
# lengths exactly (validation/manifest.d/core.yaml:90-96, 134-140).
def networkx_documented(fx, variant):
    tags = set(fx["tags"])
    if "multi_edges" in tags and variant in ("betweenness", "betweenness_normalized"):
        return True
    if variant == "betweenness_weighted_strength" and "integer_weights" in tags:
        return True
    return False
