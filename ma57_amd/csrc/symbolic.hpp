// symbolic.hpp -- CPU symbolic analysis: fill-reducing ordering, elimination
// tree, column counts, fundamental supernodes and relaxed amalgamation.
//
// Produces the same kind of result as the reference's SymbolicFactor::analyze
// (include/symla/symbolic.hpp:195-399): a permutation (perm[k] = original
// index of the k-th pivot), the elimination tree of the permuted pattern, and
// supernodes whose own columns are contiguous and lead their sorted row
// pattern.  The supernode construction keeps the reference's rules
// (postorder with children in increasing index order, the fundamental
// supernode test, bottom-up absorption of children under a size cap and a
// relative fill cap, relabeling in a postorder of the merged groups) so
// that, for the same ordering, the supernode partition is the same.  The orderings themselves are this
// project's own (SuiteSparse AMD and METIS are not available).
#pragma once

#include <cstdint>
#include <vector>

namespace symla_hip {

enum class Ordering : int { AMD = 0, Metis = 1, Natural = 2, User = 3 };

struct Supernode {
  int firstCol = 0;
  int ncols = 0;
  std::vector<int> rowPattern;  // sorted; first ncols entries are the own columns
};

struct SymbolicResult {
  int n = 0;
  std::vector<int> perm;          // final order -> original index
  std::vector<int> etreeParent;   // elimination tree in final order (-1 = root)
  std::vector<int> colCount;      // |Struct(L*j)| incl. diagonal, final order
  std::vector<Supernode> supernodes;
  int64_t nnzL = 0;               // sum of colCount
};

struct SymbolicOptions {
  int max_relax_size = 64;            // symbolic.hpp:162
  double max_relax_fill_fraction = 0.25;  // symbolic.hpp:183
};

// Full symmetric adjacency (A + A^T) without the diagonal, CSR with sorted,
// de-duplicated neighbours (ordering.hpp:44-75 semantics).
struct Graph {
  int n = 0;
  std::vector<int64_t> ptr;
  std::vector<int> adj;
};

Graph buildSymmetricGraph(int n, const int64_t* colptr, const int* rowidx);

// Minimum-degree ordering on a quotient graph with approximate external
// degrees, supervariables and element absorption.  Returns perm (position ->
// vertex).
std::vector<int> minimumDegreeOrder(const Graph& g);

// Nested dissection: level-structure vertex separators from several
// breadth-first roots, minimum degree on small leaves.  Returns perm (position -> vertex).
std::vector<int> nestedDissectionOrder(const Graph& g);


// Full pipeline.  `userPerm` is used when ordering == User.
SymbolicResult analyze(int n, const int64_t* colptr, const int* rowidx, Ordering ordering,
                       const SymbolicOptions& opts, const int* userPerm = nullptr);

// Block-diagonal union of independent analyses: system s's indices shift by
// the sizes of the systems before it.  Concatenated postordered forests stay
// postordered, so the result is the analysis of the block-diagonal matrix
// (each system analyzed on its own: orderings never mix systems).
SymbolicResult concatenateAnalyses(const std::vector<SymbolicResult>& parts);

// Elimination tree (Liu, path compression) of the graph under permutation
// `perm`; returns parent in permuted index space.
std::vector<int> eliminationTree(const Graph& g, const std::vector<int>& perm, const std::vector<int>& inv);

// Postorder of a forest: children visited in increasing index order, roots in
// increasing order (symbolic.hpp:85-105).  Returns order[k] = node visited k-th.
std::vector<int> postorderForest(const std::vector<int>& parent);

}  // namespace symla_hip
