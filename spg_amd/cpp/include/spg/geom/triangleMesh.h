// Mirror of spg::TriangleMesh (spg/include/spg/geom/triangleMesh.h:16-59) and spg::TetrahedralMesh
// (tetrahedralMesh.h:12-51): validation, the edge list in (min, max) lexicographic order and, per
// edge, the opposing vertices in face-discovery order — the order that fixes the bending stencils.
#pragma once

#include <algorithm>
#include <map>
#include <set>
#include <stdexcept>
#include <vector>

#include <spg/types.h>

namespace spg {

class TriangleMesh
{
public:
    struct OpposingVertex {
        int triangleId;
        int vertexId;
    };

    TriangleMesh(const std::vector<Vector3> &vertices, const std::vector<Vector3> &vertices0,
                 const std::vector<Int3> &faces);
    const std::vector<Vector3> &vertices() const { return m_vertices; }
    const std::vector<Vector3> &vertices0() const { return m_vertices0; }
    const std::vector<Int3> &faces() const { return m_faces; }
    const std::vector<Int2> &edges() const { return m_edges; }
    const std::vector<std::vector<OpposingVertex>> &opposingVerticesPerEdge() const { return m_opposingVerticesPerEdge; }
    // {e0, e1, opp0, opp1} for every edge with exactly two opposing vertices, as the cloth scenes of
    // apps/solverCompare.cpp:441-452 wire DiscreteBendingEnergy.
    std::vector<Int4> bendingStencils() const;

protected:
    std::vector<Vector3> m_vertices;
    std::vector<Vector3> m_vertices0;
    std::vector<Int3> m_faces;
    std::vector<Int2> m_edges;
    std::vector<std::vector<OpposingVertex>> m_opposingVerticesPerEdge;
};

class TetrahedralMesh
{
public:
    TetrahedralMesh(const std::vector<Vector3> &vertices, const std::vector<Vector3> &vertices0,
                    const std::vector<Int4> &tets);
    const std::vector<Vector3> &vertices() const { return m_vertices; }
    const std::vector<Vector3> &vertices0() const { return m_vertices0; }
    const std::vector<Int4> &tets() const { return m_tets; }
    const std::vector<Int2> &edges() const { return m_edges; }

protected:
    std::vector<Vector3> m_vertices;
    std::vector<Vector3> m_vertices0;
    std::vector<Int4> m_tets;
    std::vector<Int2> m_edges;
};

}  // namespace spg
