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
