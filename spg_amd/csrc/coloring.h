// Greedy vertex colouring over stencil cliques (coloring::colorVertices).
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

}  // namespace spgvbd
