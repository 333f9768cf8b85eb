// symbolic.cpp -- CPU symbolic analysis (see symbolic.hpp).
#include "symbolic.hpp"

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <unordered_map>
#include <vector>

namespace symla_hip {

Graph buildSymmetricGraph(int n, const int64_t* colptr, const int* rowidx) {
  // Pattern of A + A^T without the diagonal (ordering.hpp:44-75 semantics):
  // count, fill, then sort + unique per vertex.
  Graph g;
  g.n = n;
  std::vector<int64_t> cnt(n + 1, 0);
  for (int c = 0; c < n; ++c)
    for (int64_t t = colptr[c]; t < colptr[c + 1]; ++t) {
      const int r = rowidx[t];
      if (r == c) continue;
      if (r < 0 || r >= n) throw std::invalid_argument("symla: row index out of range");
      ++cnt[c + 1];
      ++cnt[r + 1];
    }
  std::vector<int64_t> ptr(n + 1, 0);
  for (int i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + cnt[i + 1];
  std::vector<int> tmp(ptr[n]);
  std::vector<int64_t> pos(ptr.begin(), ptr.end() - 1);
  for (int c = 0; c < n; ++c)
    for (int64_t t = colptr[c]; t < colptr[c + 1]; ++t) {
      const int r = rowidx[t];
      if (r == c) continue;
      tmp[pos[c]++] = r;
      tmp[pos[r]++] = c;
    }
  g.ptr.assign(n + 1, 0);
  g.adj.reserve(tmp.size());
  for (int i = 0; i < n; ++i) {
    auto b = tmp.begin() + ptr[i], e = tmp.begin() + ptr[i + 1];
    std::sort(b, e);
    auto u = std::unique(b, e);
    g.adj.insert(g.adj.end(), b, u);
    g.ptr[i + 1] = static_cast<int64_t>(g.adj.size());
  }
  return g;
}

// ---------------------------------------------------------------------------
// Minimum degree on the quotient graph.
//
// Variables carry an approximate external degree; eliminating the variable p
// of least degree turns it into an element whose variable list is the union
// of its variable neighbours and the lists of the elements it absorbs.  The
// degree of each variable i in that list is bounded by
//   |A_i| + |L_p \ i| + sum_{e in E_i, e != p} |L_e \ L_p|
// (the bound of Amestoy, Davis and Duff, SIMAX 1996).  Variables of the new
// element with identical element and variable adjacency are merged into
// supervariables and leave together.  Written from the published method.
// ---------------------------------------------------------------------------
std::vector<int> minimumDegreeOrder(const Graph& g) {
  const int n = g.n;
  std::vector<int> order;
  order.reserve(n);
  if (n == 0) return order;

  enum : char { VAR = 0, ELEM = 1, DEAD = 2 };
  std::vector<char> status(n, VAR);
  std::vector<int> nv(n, 1);
  std::vector<std::vector<int>> varAdj(n), elemAdj(n), Le(n);
  std::vector<int> degree(n, 0);
  std::vector<int> memberNext(n, -1), memberTail(n);
  for (int i = 0; i < n; ++i) {
    varAdj[i].assign(g.adj.begin() + g.ptr[i], g.adj.begin() + g.ptr[i + 1]);
    degree[i] = static_cast<int>(varAdj[i].size());
    memberTail[i] = i;
  }
  // degree buckets (doubly linked)
  std::vector<int> head(n + 1, -1), nxt(n, -1), prv(n, -1);
  auto bucketInsert = [&](int i) {
    int d = std::min(std::max(degree[i], 0), n);
    nxt[i] = head[d];
    prv[i] = -1;
    if (head[d] != -1) prv[head[d]] = i;
    head[d] = i;
  };
  auto bucketRemove = [&](int i) {
    int d = std::min(std::max(degree[i], 0), n);
    if (prv[i] != -1)
      nxt[prv[i]] = nxt[i];
    else
      head[d] = nxt[i];
    if (nxt[i] != -1) prv[nxt[i]] = prv[i];
    nxt[i] = prv[i] = -1;
  };
  for (int i = 0; i < n; ++i) bucketInsert(i);

  std::vector<int> mark(n, 0);
  int tag = 0;
  std::vector<int> wval(n, 0), wtag(n, 0);
  int wt = 0;
  int nel = 0;
  int mindeg = 0;
  std::vector<int> Lp;

  while (nel < n) {
    while (mindeg <= n && head[mindeg] == -1) ++mindeg;
    if (mindeg > n) break;  // should not happen
    const int p = head[mindeg];
    bucketRemove(p);

    // --- form element p ---
    ++tag;
    mark[p] = tag;
    Lp.clear();
    for (int i : varAdj[p])
      if (status[i] == VAR && nv[i] > 0 && mark[i] != tag) {
        mark[i] = tag;
        Lp.push_back(i);
      }
    for (int e : elemAdj[p]) {
      if (status[e] != ELEM) continue;
      for (int i : Le[e])
        if (status[i] == VAR && nv[i] > 0 && mark[i] != tag) {
          mark[i] = tag;
          Lp.push_back(i);
        }
      status[e] = DEAD;  // absorbed into p
      std::vector<int>().swap(Le[e]);
    }
    status[p] = ELEM;
    std::vector<int>().swap(varAdj[p]);
    std::vector<int>().swap(elemAdj[p]);
    for (int v = p; v != -1; v = memberNext[v]) order.push_back(v);
    nel += nv[p];

    int degLp = 0;
    for (int i : Lp) degLp += nv[i];

    // --- update adjacency of the variables in Lp ---
    for (int i : Lp) {
      bucketRemove(i);
      auto& E = elemAdj[i];
      size_t w = 0;
      for (size_t t = 0; t < E.size(); ++t)
        if (status[E[t]] == ELEM && E[t] != p) E[w++] = E[t];
      E.resize(w);
      E.push_back(p);
      auto& A = varAdj[i];
      w = 0;
      for (size_t t = 0; t < A.size(); ++t) {
        const int j = A[t];
        if (status[j] == VAR && nv[j] > 0 && mark[j] != tag) A[w++] = j;
      }
      A.resize(w);
    }

    // --- |L_e \ L_p| for elements adjacent to Lp ---
    ++wt;
    for (int i : Lp) {
      for (int e : elemAdj[i]) {
        if (e == p || wtag[e] == wt) continue;
        wtag[e] = wt;
        auto& L = Le[e];
        int s = 0;
        size_t w = 0;
        for (size_t t = 0; t < L.size(); ++t) {
          const int j = L[t];
          if (status[j] != VAR || nv[j] == 0) continue;
          L[w++] = j;
          if (mark[j] != tag) s += nv[j];
        }
        L.resize(w);
        wval[e] = s;
      }
    }

    // --- approximate degrees ---
    const int remaining = n - nel;
    for (int i : Lp) {
      int dA = 0;
      for (int j : varAdj[i]) dA += nv[j];
      long long dE = 0;
      auto& E = elemAdj[i];
      size_t w = 0;
      for (size_t t = 0; t < E.size(); ++t) {
        const int e = E[t];
        if (e == p) {
          E[w++] = e;
          continue;
        }
        if (wval[e] == 0) {
          // L_e is contained in L_p: element e is absorbed into p.
          status[e] = DEAD;
          std::vector<int>().swap(Le[e]);
          continue;
        }
        dE += wval[e];
        E[w++] = e;
      }
      E.resize(w);
      long long d = static_cast<long long>(dA) + (degLp - nv[i]) + dE;
      d = std::min<long long>(d, static_cast<long long>(degree[i]) + degLp - nv[i]);
      d = std::min<long long>(d, remaining - nv[i]);
      degree[i] = static_cast<int>(std::max<long long>(d, 0));
    }

    // --- supervariable detection among Lp ---
    if (Lp.size() > 1) {
      std::unordered_map<uint64_t, std::vector<int>> buckets;
      buckets.reserve(Lp.size() * 2);
      for (int i : Lp) {
        auto& E = elemAdj[i];
        auto& A = varAdj[i];
        std::sort(E.begin(), E.end());
        std::sort(A.begin(), A.end());
        uint64_t h = E.size() * 1000003ull + A.size();
        for (int e : E) h = h * 31 + static_cast<uint64_t>(e);
        for (int j : A) h = h * 37 + static_cast<uint64_t>(j);
        buckets[h].push_back(i);
      }
      for (auto& kv : buckets) {
        auto& lst = kv.second;
        for (size_t a = 0; a < lst.size(); ++a) {
          const int i = lst[a];
          if (nv[i] == 0) continue;
          for (size_t b = a + 1; b < lst.size(); ++b) {
            const int j = lst[b];
            if (nv[j] == 0) continue;
            if (elemAdj[i] == elemAdj[j] && varAdj[i] == varAdj[j]) {
              // merge j into i
              degree[i] -= nv[j];
              nv[i] += nv[j];
              nv[j] = 0;
              status[j] = DEAD;
              memberNext[memberTail[i]] = j;
              memberTail[i] = memberTail[j];
              std::vector<int>().swap(elemAdj[j]);
              std::vector<int>().swap(varAdj[j]);
            }
          }
        }
      }
    }

    // --- element p's list keeps the surviving principal variables ---
    auto& L = Le[p];
    L.clear();
    for (int i : Lp)
      if (nv[i] > 0) {
        L.push_back(i);
        if (degree[i] < 0) degree[i] = 0;
        bucketInsert(i);
        mindeg = std::min(mindeg, std::min(degree[i], n));
      }
  }
  if (static_cast<int>(order.size()) != n) throw std::runtime_error("symla: minimum degree ordering failed");
  return order;
}

// ---------------------------------------------------------------------------
// Nested dissection.
//
// Each connected piece above the leaf size is split by a level-structure
// vertex separator (levelSplit below) into two sides and the separator; the
// sides are dissected in turn and the separator is ordered after them.  Pieces at or
// below the leaf size, and separators, are ordered by minimum degree on their
// induced subgraph.  Order = [side A, side B, separator].
// ---------------------------------------------------------------------------
namespace {

struct NDState {
  const Graph& g;
  std::vector<int>& out;
  std::vector<int> where;   // piece id of each vertex (-1 = ordered)
  std::vector<int> mark;    // BFS scratch
  std::vector<int> local;   // global -> local index scratch
  int nextPiece = 1;
  NDState(const Graph& gg, std::vector<int>& o)
      : g(gg), out(o), where(gg.n, 0), mark(gg.n, -1), local(gg.n, -1) {}

  // Subgraph induced by the vertices V of one piece, in V's local numbering.
  Graph induced(const std::vector<int>& V) {
    Graph sub;
    sub.n = static_cast<int>(V.size());
    for (int t = 0; t < sub.n; ++t) local[V[t]] = t;
    sub.ptr.assign(sub.n + 1, 0);
    const int pid = where[V[0]];
    for (int t = 0; t < sub.n; ++t) {
      const int v = V[t];
      for (int64_t q = g.ptr[v]; q < g.ptr[v + 1]; ++q) {
        const int u = g.adj[q];
        if (where[u] == pid) sub.adj.push_back(local[u]);
      }
      sub.ptr[t + 1] = static_cast<int64_t>(sub.adj.size());
    }
    for (int v : V) local[v] = -1;
    return sub;
  }

  void orderLeaf(const std::vector<int>& V) {
    std::vector<int> o = minimumDegreeOrder(induced(V));
    for (int t : o) out.push_back(V[t]);
    for (int v : V) where[v] = -1;
  }

  // Connected components of piece `pid` among V (breadth-first).
  std::vector<std::vector<int>> components(const std::vector<int>& V, int pid) {
    std::vector<std::vector<int>> comps;
    for (int v : V) {
      if (mark[v] >= 0) continue;
      std::vector<int> q{v};
      mark[v] = 1;
      for (size_t h = 0; h < q.size(); ++h) {
        const int x = q[h];
        for (int64_t t = g.ptr[x]; t < g.ptr[x + 1]; ++t) {
          const int u = g.adj[t];
          if (where[u] == pid && mark[u] < 0) {
            mark[u] = 1;
            q.push_back(u);
          }
        }
      }
      comps.push_back(std::move(q));
    }
    for (int v : V) mark[v] = -1;
    return comps;
  }

  // Level-structure split of a connected local graph.  Breadth-first levels
  // from a root; the separator is the smallest level leaving both sides at
  // most 70% of the piece, thinned to the vertices that touch the next
  // level.  The root matters: from a corner of a box the levels are shells,
  // from the middle of a face they become planes once they span the face.
  // So several roots are tried -- a pseudo-peripheral vertex and vertices
  // spread over the farthest level seen from it and from its own farthest
  // vertex -- and the smallest separator wins (ties: the better balance).
  // part[v]: 0 / 1 = the sides, 2 = separator; empty when no level splits.
  static std::vector<int> levelSplit(const Graph& s) {
    const int n = s.n;
    std::vector<int> lev(n, -1), q;
    q.reserve(n);
    auto bfs = [&](int src) {
      std::fill(lev.begin(), lev.end(), -1);
      q.clear();
      q.push_back(src);
      lev[src] = 0;
      for (size_t h = 0; h < q.size(); ++h) {
        const int v = q[h];
        for (int64_t t = s.ptr[v]; t < s.ptr[v + 1]; ++t)
          if (lev[s.adj[t]] < 0) {
            lev[s.adj[t]] = lev[v] + 1;
            q.push_back(s.adj[t]);
          }
      }
      return lev[q.back()];
    };
    // vertices of the farthest level of the last search, up to k spread evenly
    auto farLevel = [&](int k) {
      const int ecc = lev[q.back()];
      size_t first = q.size();
      while (first > 0 && lev[q[first - 1]] == ecc) --first;
      std::vector<int> out;
      const size_t cnt = q.size() - first;
      for (int i = 0; i < k && static_cast<size_t>(i) < cnt; ++i) out.push_back(q[first + (cnt * i) / k]);
      return out;
    };
    constexpr int kRoots = 8;
    std::vector<int> roots;
    bfs(0);
    const int p1 = farLevel(1)[0];
    roots.push_back(p1);
    bfs(p1);
    const std::vector<int> f1 = farLevel(kRoots);
    roots.insert(roots.end(), f1.begin(), f1.end());
    bfs(f1[0]);
    const std::vector<int> f2 = farLevel(kRoots);
    roots.insert(roots.end(), f2.begin(), f2.end());

    int bestRoot = -1, bestL = -1, bestS = 0, bestBig = 0;
    int fbRoot = -1, fbL = -1, fbBig = n + 1;
    std::vector<int> cnt;
    for (int r : roots) {
      const int ecc = bfs(r);
      cnt.assign(ecc + 2, 0);
      for (int v = 0; v < n; ++v) ++cnt[lev[v]];
      int before = 0;
      for (int L = 0; L <= ecc; ++L) {
        const int a = before, sz = cnt[L], b = n - before - sz;
        before += sz;
        if (a == 0 || b == 0) continue;
        const int big = std::max(a, b);
        if (big < fbBig) {
          fbBig = big;
          fbL = L;
          fbRoot = r;
        }
        if (10 * static_cast<int64_t>(big) > 7 * static_cast<int64_t>(n)) continue;
        if (bestL < 0 || sz < bestS || (sz == bestS && big < bestBig)) {
          bestL = L;
          bestS = sz;
          bestBig = big;
          bestRoot = r;
        }
      }
    }
    if (bestL < 0) {
      bestL = fbL;
      bestRoot = fbRoot;
    }
    if (bestL < 0) return {};
    bfs(bestRoot);
    std::vector<int> part(n);
    for (int v = 0; v < n; ++v) {
      if (lev[v] < bestL) {
        part[v] = 0;
      } else if (lev[v] > bestL) {
        part[v] = 1;
      } else {
        bool touchesNext = false;
        for (int64_t t = s.ptr[v]; t < s.ptr[v + 1] && !touchesNext; ++t) touchesNext = lev[s.adj[t]] == bestL + 1;
        part[v] = touchesNext ? 2 : 0;
      }
    }
    return part;
  }

  void dissect(std::vector<int> V) {
    const int leafSize = 192;
    if (V.empty()) return;
    if (static_cast<int>(V.size()) <= leafSize) {
      orderLeaf(V);
      return;
    }
    const int pid = where[V[0]];
    std::vector<std::vector<int>> comps = components(V, pid);
    if (comps.size() > 1) {
      for (auto& c : comps) {
        const int np = nextPiece++;
        for (int v : c) where[v] = np;
        dissect(std::move(c));
      }
      return;
    }
    const Graph sub = induced(V);
    std::vector<int> part = levelSplit(sub);
    if (part.empty()) {
      orderLeaf(V);
      return;
    }
    std::vector<int> A, B, S;
    for (size_t t = 0; t < V.size(); ++t) (part[t] == 0 ? A : part[t] == 1 ? B : S).push_back(V[t]);
    if (A.empty() || B.empty()) {  // no useful split: one leaf
      orderLeaf(V);
      return;
    }
    const int pa = nextPiece++, pb = nextPiece++, ps = nextPiece++;
    for (int v : A) where[v] = pa;
    for (int v : B) where[v] = pb;
    for (int v : S) where[v] = ps;
    dissect(std::move(A));
    dissect(std::move(B));
    if (!S.empty()) orderLeaf(S);
  }
};

}  // namespace

std::vector<int> nestedDissectionOrder(const Graph& g) {
  std::vector<int> out;
  out.reserve(g.n);
  if (g.n == 0) return out;
  NDState st(g, out);
  std::vector<int> all(g.n);
  std::iota(all.begin(), all.end(), 0);
  st.dissect(std::move(all));
  if (static_cast<int>(out.size()) != g.n) throw std::runtime_error("symla: nested dissection failed");
  return out;
}

SymbolicResult concatenateAnalyses(const std::vector<SymbolicResult>& parts) {
  SymbolicResult r;
  for (const auto& p : parts) r.n += p.n;
  r.perm.reserve(r.n);
  r.etreeParent.reserve(r.n);
  r.colCount.reserve(r.n);
  size_t nsn = 0;
  for (const auto& p : parts) nsn += p.supernodes.size();
  r.supernodes.reserve(nsn);
  int off = 0;
  for (const auto& p : parts) {
    for (int v : p.perm) r.perm.push_back(v + off);
    for (int v : p.etreeParent) r.etreeParent.push_back(v < 0 ? -1 : v + off);
    r.colCount.insert(r.colCount.end(), p.colCount.begin(), p.colCount.end());
    for (const Supernode& s : p.supernodes) {
      Supernode t;
      t.firstCol = s.firstCol + off;
      t.ncols = s.ncols;
      t.rowPattern.reserve(s.rowPattern.size());
      for (int v : s.rowPattern) t.rowPattern.push_back(v + off);
      r.supernodes.push_back(std::move(t));
    }
    r.nnzL += p.nnzL;
    off += p.n;
  }
  return r;
}

// ---------------------------------------------------------------------------
// Elimination tree and postorder.
// ---------------------------------------------------------------------------
std::vector<int> eliminationTree(const Graph& g, const std::vector<int>& perm, const std::vector<int>& inv) {
  // Liu's algorithm with path compression (elimination_tree.hpp:78-96
  // describes the same textbook method).
  const int n = g.n;
  std::vector<int> parent(n, -1), ancestor(n, -1);
  for (int j = 0; j < n; ++j) {
    const int v = perm[j];
    for (int64_t t = g.ptr[v]; t < g.ptr[v + 1]; ++t) {
      int r = inv[g.adj[t]];
      if (r >= j) continue;
      while (ancestor[r] != -1 && ancestor[r] != j) {
        const int nx = ancestor[r];
        ancestor[r] = j;
        r = nx;
      }
      if (ancestor[r] == -1) {
        ancestor[r] = j;
        parent[r] = j;
      }
    }
  }
  return parent;
}

std::vector<int> postorderForest(const std::vector<int>& parent) {
  const int n = static_cast<int>(parent.size());
  // children lists in increasing index order (head insertion in reverse)
  std::vector<int> firstChild(n, -1), nextSib(n, -1);
  for (int j = n - 1; j >= 0; --j) {
    const int p = parent[j];
    if (p != -1) {
      nextSib[j] = firstChild[p];
      firstChild[p] = j;
    }
  }
  std::vector<int> order;
  order.reserve(n);
  std::vector<int> stack;
  std::vector<int> it(n, -2);
  for (int r = 0; r < n; ++r) {
    if (parent[r] != -1) continue;
    stack.push_back(r);
    it[r] = firstChild[r];
    while (!stack.empty()) {
      const int v = stack.back();
      const int c = it[v];
      if (c != -1) {
        it[v] = nextSib[c];
        stack.push_back(c);
        it[c] = firstChild[c];
      } else {
        order.push_back(v);
        stack.pop_back();
      }
    }
  }
  return order;
}

// ---------------------------------------------------------------------------
// Pipeline (symbolic.hpp:195-399 of the reference, with column counts by
// row-subtree traversal and supernode patterns built supernode-wise instead
// of materializing every column's pattern).
// ---------------------------------------------------------------------------

SymbolicResult analyze(int n, const int64_t* colptr, const int* rowidx, Ordering ordering,
                       const SymbolicOptions& opts, const int* userPerm) {
  SymbolicResult sr;
  sr.n = n;
  if (n == 0) return sr;
  Graph g = buildSymmetricGraph(n, colptr, rowidx);

  // --- step 1: fill-reducing order ---
  std::vector<int> perm0;
  switch (ordering) {
    case Ordering::Natural:
      perm0.resize(n);
      std::iota(perm0.begin(), perm0.end(), 0);
      break;
    case Ordering::AMD:
      perm0 = minimumDegreeOrder(g);
      break;
    case Ordering::Metis:
      perm0 = nestedDissectionOrder(g);
      break;
    case Ordering::User: {
      if (!userPerm) throw std::invalid_argument("symla: user ordering requires a permutation");
      perm0.assign(userPerm, userPerm + n);
      std::vector<char> seen(n, 0);
      for (int v : perm0) {
        if (v < 0 || v >= n || seen[v]) throw std::invalid_argument("symla: user ordering is not a permutation");
        seen[v] = 1;
      }
      break;
    }
  }
  std::vector<int> inv0(n);
  for (int k = 0; k < n; ++k) inv0[perm0[k]] = k;

  // --- step 3: etree + postorder (symbolic.hpp:212-216) ---
  std::vector<int> tree0 = eliminationTree(g, perm0, inv0);
  std::vector<int> post = postorderForest(tree0);
  std::vector<int> finalPerm(n);
  for (int k = 0; k < n; ++k) finalPerm[k] = perm0[post[k]];
  std::vector<int> inv1(n);
  for (int k = 0; k < n; ++k) inv1[finalPerm[k]] = k;

  // --- step 4: etree on the postordered pattern + column counts ---
  std::vector<int> parent = eliminationTree(g, finalPerm, inv1);
  std::vector<int> colCount(n, 1);
  {
    // row-subtree traversal: row i of L is the union of etree paths from
    // each j in Adj(i), j < i, up to i.
    std::vector<int> flag(n, -1);
    for (int i = 0; i < n; ++i) {
      flag[i] = i;
      const int v = finalPerm[i];
      for (int64_t t = g.ptr[v]; t < g.ptr[v + 1]; ++t) {
        int j = inv1[g.adj[t]];
        if (j >= i) continue;
        while (flag[j] != i) {
          flag[j] = i;
          ++colCount[j];
          j = parent[j];
        }
      }
    }
  }

  // --- step 5: chains of the postordered etree (symbolic.hpp:229-247) ---
  // A column continues the chain of the column before it when it is that
  // column's etree parent and its count is one less; a chain is a
  // fundamental supernode.  chainOf[c] = chain id, chains numbered upwards.
  std::vector<int> chainOf(n);
  std::vector<int> chainLo, chainHi;  // [lo, hi] stage-1 columns of each chain
  for (int c = 0; c < n; ++c) {
    const bool extends = c > 0 && parent[c - 1] == c && colCount[c - 1] == colCount[c] + 1;
    if (!extends) {
      chainLo.push_back(c);
      chainHi.push_back(c);
    } else {
      chainHi.back() = c;
    }
    chainOf[c] = static_cast<int>(chainLo.size()) - 1;
  }
  const int nch = static_cast<int>(chainLo.size());
  // chain tree: the chain above is the one holding the etree parent of the top column
  std::vector<int> chainUp(nch, -1);
  for (int f = 0; f < nch; ++f) {
    const int p = parent[chainHi[f]];
    chainUp[f] = p < 0 ? -1 : chainOf[p];
  }
  // children of each chain in increasing id (CSR; counting sort keeps the order)
  std::vector<int> kidPtr(nch + 1, 0), kids;
  for (int f = 0; f < nch; ++f)
    if (chainUp[f] >= 0) ++kidPtr[chainUp[f] + 1];
  for (int f = 0; f < nch; ++f) kidPtr[f + 1] += kidPtr[f];
  kids.resize(kidPtr[nch]);
  {
    std::vector<int> fill(kidPtr.begin(), kidPtr.end() - 1);
    for (int f = 0; f < nch; ++f)
      if (chainUp[f] >= 0) kids[fill[chainUp[f]]++] = f;
  }
  // Rows of each chain below its own columns (stage-1 indices, sorted): the
  // A entries below the chain's columns plus the children's rows above
  // their own chains (symbolic elimination chain by chain).  The count must
  // equal colCount of the chain's first column.
  std::vector<std::vector<int>> below(nch);
  {
    std::vector<int> seen(n, -1);
    for (int f = 0; f < nch; ++f) {
      const int lo = chainLo[f], hi = chainHi[f];
      auto& rows = below[f];
      auto take = [&](int r) {
        if (r > hi && seen[r] != f) {
          seen[r] = f;
          rows.push_back(r);
        }
      };
      for (int c = lo; c <= hi; ++c) {
        const int v = finalPerm[c];
        for (int64_t t = g.ptr[v]; t < g.ptr[v + 1]; ++t) take(inv1[g.adj[t]]);
      }
      for (int t = kidPtr[f]; t < kidPtr[f + 1]; ++t)
        for (int r : below[kids[t]]) take(r);
      std::sort(rows.begin(), rows.end());
      if (static_cast<int>(rows.size()) + (hi - lo + 1) != colCount[lo])
        throw std::runtime_error("symla: supernode pattern does not match column count");
    }
  }

  // --- step 6: relaxed amalgamation (symbolic.hpp:269-329 rules) ---
  // Chains are visited upwards; each tries to absorb its children in
  // increasing id, the child bringing every chain it absorbed before.  The
  // test needs only sizes: a child group's rows lie in its own columns plus
  // the parent group's rows (etree nesting), and its columns are disjoint
  // from the parent group's rows, so absorbing adds exactly the child
  // group's columns to both the column count and the row count of the
  // parent group.  Storage of a k-column, m-row block: k m - k (k - 1) / 2.
  auto storage = [](int64_t k, int64_t m) { return k * m - k * (k - 1) / 2; };
  std::vector<int> gCols(nch), gRows(nch);  // group totals, valid at group tops
  std::vector<char> absorbed(nch, 0);       // chain f was absorbed by chainUp[f]
  for (int f = 0; f < nch; ++f) {
    gCols[f] = chainHi[f] - chainLo[f] + 1;
    gRows[f] = colCount[chainLo[f]];
    for (int t = kidPtr[f]; t < kidPtr[f + 1]; ++t) {
      const int c = kids[t];
      const int cols = gCols[f] + gCols[c];
      if (cols > opts.max_relax_size) continue;
      const int rows = gRows[f] + gCols[c];
      const int64_t joint = storage(cols, rows);
      const int64_t apart = storage(gCols[f], gRows[f]) + storage(gCols[c], gRows[c]);
      if (static_cast<double>(joint - apart) / static_cast<double>(joint) > opts.max_relax_fill_fraction) continue;
      absorbed[c] = 1;
      gCols[f] = cols;
      gRows[f] = rows;
    }
  }
  // group top of every chain: parents have larger ids, so sweep downwards
  std::vector<int> top(nch);
  for (int f = nch - 1; f >= 0; --f) top[f] = absorbed[f] ? top[chainUp[f]] : f;

  // --- step 6': relabel so every group is contiguous (symbolic.hpp:335-366 rules) ---
  // Groups in postorder of the group forest (children and roots by
  // increasing top id); inside a group, its chains in postorder of the
  // absorbed edges (children by increasing id, the top last).
  std::vector<int> tops;
  std::vector<int> slot(nch, -1);
  for (int f = 0; f < nch; ++f)
    if (!absorbed[f]) {
      slot[f] = static_cast<int>(tops.size());
      tops.push_back(f);
    }
  const int ngroups = static_cast<int>(tops.size());
  std::vector<int> groupUp(ngroups, -1);
  for (int q = 0; q < ngroups; ++q) {
    const int u = chainUp[tops[q]];
    groupUp[q] = u < 0 ? -1 : slot[top[u]];
  }
  const std::vector<int> groupOrder = postorderForest(groupUp);
  std::vector<int> inv2(n);  // stage-1 index -> final index
  std::vector<int> groupTopOf;
  groupTopOf.reserve(ngroups);
  sr.supernodes.reserve(ngroups);
  {
    int next = 0;
    std::vector<std::pair<int, int>> dfs;  // (chain, next child slot)
    for (int q : groupOrder) {
      const int f0 = tops[q];
      Supernode sn;
      sn.firstCol = next;
      sn.ncols = gCols[f0];
      dfs.assign(1, {f0, kidPtr[f0]});
      while (!dfs.empty()) {
        auto& [f, t] = dfs.back();
        while (t < kidPtr[f + 1] && !absorbed[kids[t]]) ++t;
        if (t < kidPtr[f + 1]) {
          const int c = kids[t++];
          dfs.push_back({c, kidPtr[c]});
          continue;
        }
        for (int c = chainLo[f]; c <= chainHi[f]; ++c) inv2[c] = next++;
        dfs.pop_back();
      }
      sr.supernodes.push_back(std::move(sn));
      groupTopOf.push_back(f0);
    }
  }

  // --- final permutation, etree, counts and supernode row patterns ---
  // A group's rows are its own (now contiguous) columns followed by the
  // rows of its top chain above that chain.
  sr.perm.resize(n);
  for (int c = 0; c < n; ++c) sr.perm[inv2[c]] = finalPerm[c];
  sr.etreeParent.assign(n, -1);
  sr.colCount.assign(n, 0);
  for (int j = 0; j < n; ++j) {
    sr.etreeParent[inv2[j]] = parent[j] == -1 ? -1 : inv2[parent[j]];
    sr.colCount[inv2[j]] = colCount[j];
  }
  for (size_t s = 0; s < sr.supernodes.size(); ++s) {
    auto& sn = sr.supernodes[s];
    const auto& up = below[groupTopOf[s]];
    std::vector<int> rows(sn.ncols);
    std::iota(rows.begin(), rows.end(), sn.firstCol);
    const size_t own = rows.size();
    for (int r : up) rows.push_back(inv2[r]);
    std::sort(rows.begin() + own, rows.end());
    if (own < rows.size() && rows[own] < sn.firstCol + sn.ncols)
      throw std::runtime_error("symla: supernode columns are not contiguous");
    sn.rowPattern.swap(rows);
  }
  sr.nnzL = 0;
  for (int c : colCount) sr.nnzL += c;
  return sr;
}

}  // namespace symla_hip
