// spg::io::loadObj (spg/include/spg/geom/io/obj.h, source/geom/io/obj.cpp:10-35): the vertices of a Wavefront
// OBJ file and the triangles of its first shape, as the reference reads them through tinyobjloader.
#pragma once

#include <string>
#include <tuple>
#include <vector>

#include <spg/types.h>

namespace spg::io {
std::tuple<std::vector<Vector3>, std::vector<Int3>> loadObj(const std::string &filePath);
}  // namespace spg::io
