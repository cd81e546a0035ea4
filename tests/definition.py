"""Exact-rational node and edge betweenness straight from the definition.

Independent of Brandes' algorithm: all-pairs shortest distances by
Floyd-Warshall over ``Fraction`` lengths, shortest-path counts by dynamic
programming over each source's distance order (parallel arcs counted
separately, loops never on a shortest path), then
b(v) = sum over ordered s != v != t of sigma_st(v) / sigma_st with
sigma_st(v) = sigma_sv * sigma_vt when d(s,v) + d(v,t) = d(s,t). Undirected
graphs count each unordered pair once; normalisation divides by (n-1)(n-2)
(directed) or (n-1)(n-2)/2. Pure Python; for graphs of a few dozen nodes.
"""
from fractions import Fraction


def _arcs(n, src, tgt, lengths, directed):
    arcs = []
    for e in range(len(src)):
        s, t = int(src[e]), int(tgt[e])
        ln = lengths[e]
        arcs.append((s, t, ln))
        if not directed and s != t:
            arcs.append((t, s, ln))
    return arcs


def betweenness(n, src, tgt, weights=None, directed=False, weighted=False,
                strength=False, normalized=False):
    if weighted:
        lengths = [Fraction(1) / Fraction(w) if strength else Fraction(w) for w in weights]
    else:
        lengths = [Fraction(1)] * len(src)
    arcs = _arcs(n, src, tgt, lengths, directed)
    INF = None
    d = [[INF] * n for _ in range(n)]
    for v in range(n):
        d[v][v] = Fraction(0)
    for s, t, ln in arcs:
        if s != t and (d[s][t] is None or ln < d[s][t]):
            d[s][t] = ln
    for k in range(n):
        dk = d[k]
        for i in range(n):
            dik = d[i][k]
            if dik is None:
                continue
            di = d[i]
            for j in range(n):
                if dk[j] is None:
                    continue
                c = dik + dk[j]
                if di[j] is None or c < di[j]:
                    di[j] = c
    # sigma[s][t]: number of shortest s-t paths (arc multiplicity counted).
    sigma = [[0] * n for _ in range(n)]
    inarcs = [[] for _ in range(n)]
    for s, t, ln in arcs:
        inarcs[t].append((s, ln))
    for s in range(n):
        order = sorted((v for v in range(n) if d[s][v] is not None), key=lambda v: d[s][v])
        sigma[s][s] = 1
        for t in order:
            if t == s:
                continue
            tot = 0
            for u, ln in inarcs[t]:
                if d[s][u] is not None and d[s][u] + ln == d[s][t]:
                    tot += sigma[s][u]
            sigma[s][t] = tot
    bc = [Fraction(0)] * n
    for s in range(n):
        for t in range(n):
            if s == t or d[s][t] is None or sigma[s][t] == 0:
                continue
            for v in range(n):
                if v == s or v == t or d[s][v] is None or d[v][t] is None:
                    continue
                if d[s][v] + d[v][t] == d[s][t]:
                    bc[v] += Fraction(sigma[s][v] * sigma[v][t], sigma[s][t])
    if not directed:
        bc = [x / 2 for x in bc]
    if normalized and n > 2:
        pairs = (n - 1) * (n - 2)
        bc = [x / pairs if directed else 2 * x / pairs for x in bc]
    return [float(x) for x in bc]


def edge_betweenness(n, src, tgt, weights=None, directed=False, weighted=False,
                     strength=False, normalized=False):
    """Per input edge e (as arcs u->v, both directions when undirected):
    sum over ordered s != t of sigma_su * sigma_vt / sigma_st for each arc
    of e with d(s,u) + len + d(v,t) = d(s,t); parallel edges are separate
    edges. Undirected graphs count each unordered pair once; normalisation
    divides by n(n-1) (directed) or n(n-1)/2."""
    if weighted:
        lengths = [Fraction(1) / Fraction(w) if strength else Fraction(w) for w in weights]
    else:
        lengths = [Fraction(1)] * len(src)
    arcs = _arcs(n, src, tgt, lengths, directed)
    d = [[None] * n for _ in range(n)]
    for v in range(n):
        d[v][v] = Fraction(0)
    for s, t, ln in arcs:
        if s != t and (d[s][t] is None or ln < d[s][t]):
            d[s][t] = ln
    for k in range(n):
        dk = d[k]
        for i in range(n):
            dik = d[i][k]
            if dik is None:
                continue
            di = d[i]
            for j in range(n):
                if dk[j] is None:
                    continue
                c = dik + dk[j]
                if di[j] is None or c < di[j]:
                    di[j] = c
    sigma = [[0] * n for _ in range(n)]
    inarcs = [[] for _ in range(n)]
    for s, t, ln in arcs:
        inarcs[t].append((s, ln))
    for s in range(n):
        order = sorted((v for v in range(n) if d[s][v] is not None), key=lambda v: d[s][v])
        sigma[s][s] = 1
        for t in order:
            if t == s:
                continue
            sigma[s][t] = sum(sigma[s][u] for u, ln in inarcs[t]
                              if d[s][u] is not None and d[s][u] + ln == d[s][t])
    eb = []
    for e in range(len(src)):
        u0, v0, ln = int(src[e]), int(tgt[e]), lengths[e]
        dirs = [(u0, v0)] if directed or u0 == v0 else [(u0, v0), (v0, u0)]
        tot = Fraction(0)
        for u, v in dirs:
            if u == v:
                continue
            for s in range(n):
                if d[s][u] is None:
                    continue
                for t in range(n):
                    if s == t or d[s][t] is None or d[v][t] is None or sigma[s][t] == 0:
                        continue
                    if d[s][u] + ln + d[v][t] == d[s][t]:
                        tot += Fraction(sigma[s][u] * sigma[v][t], sigma[s][t])
        eb.append(tot)
    if not directed:
        eb = [x / 2 for x in eb]
    if normalized:
        if n < 2:
            return [float("nan")] * len(eb)
        pairs = n * (n - 1)
        eb = [x / pairs if directed else 2 * x / pairs for x in eb]
    return [float(x) for x in eb]


def subset_raw(n, src, tgt, weights=None, directed=False, weighted=False, strength=False,
               sources=None, targets=None, endpoints=False, collapse=False):
    """Unscaled betweenness for sources S (a list; repeats count) and
    targets T (a set), from the definition: node v gets, for s in S and
    t in T with s != t reachable, sigma_st(v) / sigma_st (v not s, t), plus
    1 for s and for t with endpoints; an edge gets sigma_su sigma_vt /
    sigma_st for each of its arcs u -> v on such a path. With collapse,
    parallel arcs count once (node sequences) with the least length, and
    an adjacent pair's edge value is shared equally by its parallel edges
    of that length (loops and longer parallels get 0). Ordered pairs; the
    caller scales."""
    m = len(src)
    if weighted:
        lengths = [Fraction(1) / Fraction(w) if strength else Fraction(w) for w in weights]
    else:
        lengths = [Fraction(1)] * m
    S = list(range(n)) if sources is None else [int(s) for s in sources]
    T = set(range(n)) if targets is None else {int(t) for t in targets}
    arcs = []  # (u, v, length, edges sharing the arc's value)
    if collapse:
        best = {}
        for e in range(m):
            a, b = int(src[e]), int(tgt[e])
            for u, v in ([(a, b)] if directed or a == b else [(a, b), (b, a)]):
                if (u, v) not in best or lengths[e] < best[(u, v)]:
                    best[(u, v)] = lengths[e]
        for (u, v), ln in best.items():
            key = (u, v) if directed or u <= v else (v, u)
            tied = [e for e in range(m)
                    if ((int(src[e]), int(tgt[e])) == key or
                        (not directed and (int(tgt[e]), int(src[e])) == key))
                    and lengths[e] == ln]
            arcs.append((u, v, ln, tied))
    else:
        for e in range(m):
            a, b = int(src[e]), int(tgt[e])
            for u, v in ([(a, b)] if directed or a == b else [(a, b), (b, a)]):
                arcs.append((u, v, lengths[e], [e]))
    d = [[None] * n for _ in range(n)]
    for v in range(n):
        d[v][v] = Fraction(0)
    for u, v, ln, _ in arcs:
        if u != v and (d[u][v] is None or ln < d[u][v]):
            d[u][v] = ln
    for k in range(n):
        for i in range(n):
            if d[i][k] is None:
                continue
            for j in range(n):
                if d[k][j] is None:
                    continue
                c = d[i][k] + d[k][j]
                if d[i][j] is None or c < d[i][j]:
                    d[i][j] = c
    sigma = [[0] * n for _ in range(n)]
    inarcs = [[] for _ in range(n)]
    for u, v, ln, _ in arcs:
        inarcs[v].append((u, ln))
    for s in range(n):
        order = sorted((v for v in range(n) if d[s][v] is not None), key=lambda v: d[s][v])
        sigma[s][s] = 1
        for t in order:
            if t != s:
                sigma[s][t] = sum(sigma[s][u] for u, ln in inarcs[t]
                                  if d[s][u] is not None and d[s][u] + ln == d[s][t])
    nodes = [Fraction(0)] * n
    edges = [Fraction(0)] * m
    for s in S:
        for t in T:
            if s == t or d[s][t] is None or sigma[s][t] == 0:
                continue
            if endpoints:
                nodes[s] += 1
                nodes[t] += 1
            for v in range(n):
                if v in (s, t) or d[s][v] is None or d[v][t] is None:
                    continue
                if d[s][v] + d[v][t] == d[s][t]:
                    nodes[v] += Fraction(sigma[s][v] * sigma[v][t], sigma[s][t])
            for u, v, ln, shared in arcs:
                if u == v or d[s][u] is None or d[v][t] is None:
                    continue
                if d[s][u] + ln + d[v][t] == d[s][t]:
                    share = Fraction(sigma[s][u] * sigma[v][t], sigma[s][t]) / len(shared)
                    for e in shared:
                        edges[e] += share
    return [float(x) for x in nodes], [float(x) for x in edges]
