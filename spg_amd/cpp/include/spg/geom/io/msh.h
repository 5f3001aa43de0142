// spg::io::loadMsh (spg/include/spg/geom/io/msh.h, source/geom/io/msh.cpp:10-41): the nodes and 4-node
// tetrahedra of a Gmsh MSH 4.1 file, as the reference reads them through MshIO.
#pragma once

#include <string>
#include <tuple>
#include <vector>

#include <spg/types.h>

namespace spg::io {
std::tuple<std::vector<Vector3>, std::vector<Int4>> loadMsh(const std::string &filePath);
}  // namespace spg::io
