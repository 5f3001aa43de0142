// Greedy vertex colouring over stencil cliques (coloring::colorVertices), greedy stencil colouring
// (coloring::colorStencils) and the level order of serial Gauss-Seidel over stencils.
#pragma once

#include <cstdint>
#include <vector>

namespace spgvbd {

struct FlatStencilSet {
    const int32_t *entries;
    int64_t length;
    int32_t stencilSize;
};

// Reference: spg/source/utils/graphColoring.cpp:72-84 (a clique per stencil of every set) and
// Graph::greedyColoring (graphColoring.cpp:21-49): node 0 takes colour 0; every later node, in
// index order, takes the smallest colour not held by an already coloured neighbour. The adjacency
// is built as a CSR and the "used" flags are cleared per node over the touched colours only, so
// the cost is O(V + E) instead of the reference's O(V^2) fill; the colours are the same.
// Throws std::runtime_error("Trying to add edge to graph with invalid node indices") like
// Graph::addEdge (graphColoring.cpp:12-18) on an out-of-range index.
std::vector<int32_t> colorVertices(int64_t numVertices, const std::vector<FlatStencilSet> &sets);

// Reference: coloring::colorStencils (spg/source/utils/graphColoring.cpp:51-70) and
// Graph::greedyColoring: a node per stencil, an edge between two stencils that share a vertex; node 0
// takes colour 0 and every later stencil, in index order, the smallest colour not held by an earlier
// stencil that shares a vertex with it. Built from per-vertex stencil lists in O(sum of list
// lengths^2 / n) instead of the reference's O(n^2) fill; the colours are the same. No stencils: no
// colours (the reference indexes node 0 of an empty graph there). Vertex indices must be >= 0.
std::vector<int32_t> colorStencils(const int32_t *flat, int64_t nStencils, int32_t stencilSize);

// Levels of the serial order (XPBD with useParallelGaussSeidel = false projects stencil 0, 1, 2, ...
// one after the other, xpbd.cpp:99-104): level(s) = 1 + the largest level of an earlier stencil that
// shares a vertex with s (0 when there is none). Running the levels in increasing order, the stencils
// of one level in parallel, performs every pair of vertex-sharing projections in index order, so it
// reproduces the serial result exactly.
std::vector<int32_t> serialLevels(const int32_t *flat, int64_t nStencils, int32_t stencilSize);

}  // namespace spgvbd
