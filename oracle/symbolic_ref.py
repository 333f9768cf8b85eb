"""Pure-Python restatement of the reference's supernode construction.

TEST INFRASTRUCTURE ONLY.  Follows include/symla/symbolic.hpp:195-399 and
include/symla/elimination_tree.hpp:61-138 line for line in behaviour (not in
data structures), for small matrices, given the fill-reducing order.  The
reference's own AMD / METIS calls (ordering.hpp:99-146) are third-party and
absent here, so the checks run with the Natural ordering or a given
permutation; the product's C++ analysis must reproduce this result exactly.
"""
from __future__ import annotations

import numpy as np


def sym_pattern_nodiag(n, colptr, rowidx):
    """ordering.hpp:44-75: neighbours of every vertex in A + A^T, no diagonal."""
    nb = [set() for _ in range(n)]
    for c in range(n):
        for t in range(colptr[c], colptr[c + 1]):
            r = int(rowidx[t])
            if r != c:
                nb[c].add(r)
                nb[r].add(c)
    return [sorted(s) for s in nb]


def permute(nb, perm):
    n = len(perm)
    inv = np.empty(n, dtype=np.int64)
    inv[np.asarray(perm)] = np.arange(n)
    out = [None] * n
    for v in range(n):
        out[inv[v]] = sorted(int(inv[u]) for u in nb[v])
    return out


def etree_and_patterns(nbp):
    """elimination_tree.hpp:78-131 on an already-permuted neighbour list."""
    n = len(nbp)
    parent = [-1] * n
    ancestor = [-1] * n
    for j in range(n):
        for i in nbp[j]:
            if i >= j:
                continue
            r = i
            while ancestor[r] != -1 and ancestor[r] != j:
                t = ancestor[r]
                ancestor[r] = j
                r = t
            if ancestor[r] == -1:
                ancestor[r] = j
                parent[r] = j
    children = [[] for _ in range(n)]
    for j in range(n):
        if parent[j] != -1:
            children[parent[j]].append(j)
    pats = []
    for j in range(n):
        pat = {j} | {i for i in nbp[j] if i > j}
        for c in children[j]:
            pat |= set(pats[c][1:])
        pats.append(sorted(pat))
    return parent, [len(p) for p in pats], pats


def postorder(parent):
    """symbolic.hpp:85-105 (children in increasing index order)."""
    n = len(parent)
    children = [[] for _ in range(n)]
    roots = []
    for j in range(n):
        if parent[j] == -1:
            roots.append(j)
        else:
            children[parent[j]].append(j)
    order = []

    def visit(v):
        stack = [(v, 0)]
        while stack:
            node, i = stack.pop()
            if i < len(children[node]):
                stack.append((node, i + 1))
                stack.append((children[node][i], 0))
            else:
                order.append(node)

    for r in roots:
        visit(r)
    return order


def analyze(n, colptr, rowidx, perm0, max_relax_size=64, max_relax_fill_fraction=0.25):
    """SymbolicFactor::analyze (symbolic.hpp:195-399) for a given first-stage order perm0.

    Returns (perm, etree_parent, supernodes) with supernodes as (firstCol, ncols, rowPattern)."""
    nb = sym_pattern_nodiag(n, colptr, rowidx)
    nb0 = permute(nb, perm0)
    tree0, _, _ = etree_and_patterns(nb0)
    post = postorder(tree0)
    final_perm = [int(perm0[post[k]]) for k in range(n)]
    nb1 = permute(nb, final_perm)
    parent, colcount, pats = etree_and_patterns(nb1)

    def fundamental(par, cc, pp):
        out = []
        j = 0
        while j < n:
            start = j
            while j + 1 < n and par[j] == j + 1 and cc[j] == cc[j + 1] + 1:
                j += 1
            out.append([start, j - start + 1, list(pp[start])])
            j += 1
        return out

    funda = fundamental(parent, colcount, pats)
    nf = len(funda)
    col_to_f = [0] * n
    for i, (f, k, _) in enumerate(funda):
        for c in range(f, f + k):
            col_to_f[c] = i
    parent_f = [-1] * nf
    children_f = [[] for _ in range(nf)]
    for i, (f, k, _) in enumerate(funda):
        p = parent[f + k - 1]
        if p != -1:
            parent_f[i] = col_to_f[p]
            children_f[col_to_f[p]].append(i)
    dsu = list(range(nf))

    def find(x):
        while dsu[x] != x:
            dsu[x] = dsu[dsu[x]]
            x = dsu[x]
        return x

    agg = [[funda[i][1], list(funda[i][2]), [i]] for i in range(nf)]

    def storage(k, m):
        return k * m - k * (k - 1) // 2

    for i in range(nf):
        for c in children_f[i]:
            cr = find(c)
            if cr == i:
                continue
            mn = agg[i][0] + agg[cr][0]
            if mn > max_relax_size:
                continue
            merged = sorted(set(agg[i][1]) | set(agg[cr][1]))
            ms = storage(mn, len(merged))
            os_ = storage(agg[i][0], len(agg[i][1])) + storage(agg[cr][0], len(agg[cr][1]))
            if (ms - os_) / ms > max_relax_fill_fraction:
                continue
            dsu[cr] = i
            agg[i][0] = mn
            agg[i][1] = merged
            agg[i][2] = agg[i][2][:-1] + agg[cr][2] + agg[i][2][-1:]
    final_roots = [i for i in range(nf) if find(i) == i]
    compact = {r: k for k, r in enumerate(final_roots)}
    fpar = [(-1 if parent_f[r] == -1 else compact[find(parent_f[r])]) for r in final_roots]
    gpost = postorder(fpar)
    sns = []
    new_order = []
    for gk in gpost:
        r = final_roots[gk]
        first = len(new_order)
        for fi in agg[r][2]:
            f, k, _ = funda[fi]
            new_order.extend(range(f, f + k))
        sns.append([first, agg[r][0]])
    combined = [final_perm[new_order[k]] for k in range(n)]
    nb2 = permute(nb, combined)
    parent2, _, pats2 = etree_and_patterns(nb2)
    out = []
    for first, k in sns:
        pat = set()
        for c in range(first, first + k):
            pat |= set(pats2[c])
        out.append((first, k, sorted(pat)))
    return combined, parent2, out
