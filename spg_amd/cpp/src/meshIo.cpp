// Mesh readers of spg::io (source/geom/io/obj.cpp:10-35, msh.cpp:10-41). The reference delegates parsing to
// tinyobjloader and MshIO, two submodules absent from its repository; these are parsers of our own for the
// parts of the two formats the reference reads:
//   OBJ: "v x y z" vertices (global numbering) and the faces of the first shape (a new shape starts at an
//        "o" or "g" line once faces were read), face corners "i", "i/t", "i//n" or "i/t/n", 1-based or
//        negative (relative to the vertices read so far); polygons are split into a triangle fan, as
//        tinyobjloader's triangulation does for convex faces.
//   MSH 4.1, ASCII: the $Nodes entity blocks in file order (node tags are not used: node k of the file is
//        vertex k) and the 4-node tetrahedron blocks (element type 4) of $Elements, node tags - 1.
// Errors (missing file, malformed lines, unsupported MSH version or binary files) throw std::runtime_error.
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

#include <spg/geom/io/msh.h>
#include <spg/geom/io/obj.h>

namespace spg::io {

namespace {

int objIndex(const std::string &corner, int nVertices)
{
    const std::string head = corner.substr(0, corner.find('/'));
    size_t used = 0;
    int k = 0;
    try {
        k = std::stoi(head, &used);
    } catch (const std::exception &) {
        throw std::runtime_error("OBJ face corner is not an index: " + corner);
    }
    if (used != head.size() || k == 0) throw std::runtime_error("OBJ face corner is not an index: " + corner);
    const int idx = k > 0 ? k - 1 : nVertices + k;
    if (idx < 0 || idx >= nVertices) throw std::runtime_error("OBJ face index out of range: " + corner);
    return idx;
}

}  // namespace

std::tuple<std::vector<Vector3>, std::vector<Int3>> loadObj(const std::string &filePath)
{
    std::ifstream in(filePath);
    if (!in) throw std::runtime_error("cannot open OBJ file " + filePath);
    std::vector<Vector3> vertices;
    std::vector<Int3> faces;
    bool firstShapeDone = false;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag) || tag[0] == '#') continue;
        if (tag == "v") {
            Real x, y, z;
            if (!(ls >> x >> y >> z)) throw std::runtime_error("malformed OBJ vertex: " + line);
            vertices.push_back({x, y, z});
        } else if (tag == "o" || tag == "g") {
            if (!faces.empty()) firstShapeDone = true;
        } else if (tag == "f") {
            std::vector<int> corners;
            std::string c;
            while (ls >> c) corners.push_back(objIndex(c, static_cast<int>(vertices.size())));
            if (corners.size() < 3) throw std::runtime_error("OBJ face with fewer than 3 corners: " + line);
            if (firstShapeDone) continue;
            for (size_t k = 1; k + 1 < corners.size(); ++k) faces.push_back({corners[0], corners[k], corners[k + 1]});
        }
    }
    return {std::move(vertices), std::move(faces)};
}

std::tuple<std::vector<Vector3>, std::vector<Int4>> loadMsh(const std::string &filePath)
{
    std::ifstream in(filePath);
    if (!in) throw std::runtime_error("cannot open MSH file " + filePath);
    std::vector<Vector3> vertices;
    std::vector<Int4> tets;
    std::string section;
    bool formatOk = false;
    while (in >> section) {
        if (section == "$MeshFormat") {
            std::string version;
            int fileType = 0, dataSize = 0;
            if (!(in >> version >> fileType >> dataSize)) throw std::runtime_error("malformed $MeshFormat");
            if (version.rfind("4.1", 0) != 0) throw std::runtime_error("only MSH 4.1 files are read, not " + version);
            if (fileType != 0) throw std::runtime_error("only ASCII MSH files are read");
            formatOk = true;
        } else if (section == "$Nodes") {
            if (!formatOk) throw std::runtime_error("MSH $Nodes before $MeshFormat");
            size_t nBlocks = 0, nNodes = 0, minTag = 0, maxTag = 0;
            if (!(in >> nBlocks >> nNodes >> minTag >> maxTag)) throw std::runtime_error("malformed $Nodes");
            vertices.reserve(nNodes);
            for (size_t b = 0; b < nBlocks; ++b) {
                int dim = 0, entity = 0, parametric = 0;
                size_t n = 0;
                if (!(in >> dim >> entity >> parametric >> n)) throw std::runtime_error("malformed node block");
                if (parametric) throw std::runtime_error("parametric MSH nodes are not read");
                for (size_t k = 0; k < n; ++k) {
                    size_t tag;
                    if (!(in >> tag)) throw std::runtime_error("malformed node tags");
                }
                for (size_t k = 0; k < n; ++k) {
                    Real x, y, z;
                    if (!(in >> x >> y >> z)) throw std::runtime_error("malformed node coordinates");
                    vertices.push_back({x, y, z});
                }
            }
        } else if (section == "$Elements") {
            if (!formatOk) throw std::runtime_error("MSH $Elements before $MeshFormat");
            size_t nBlocks = 0, nElems = 0, minTag = 0, maxTag = 0;
            if (!(in >> nBlocks >> nElems >> minTag >> maxTag)) throw std::runtime_error("malformed $Elements");
            for (size_t b = 0; b < nBlocks; ++b) {
                int dim = 0, entity = 0, type = 0;
                size_t n = 0;
                if (!(in >> dim >> entity >> type >> n)) throw std::runtime_error("malformed element block");
                // nodes per element: first-order point, line, triangle, quadrangle, tetrahedron, hexahedron,
                // prism, pyramid (only the tetrahedra are kept)
                int per = 0;
                switch (type) {
                case 15: per = 1; break;
                case 1: per = 2; break;
                case 2: per = 3; break;
                case 3: per = 4; break;
                case 4: per = 4; break;
                case 5: per = 8; break;
                case 6: per = 6; break;
                case 7: per = 5; break;
                default: throw std::runtime_error("unsupported MSH element type " + std::to_string(type));
                }
                for (size_t k = 0; k < n; ++k) {
                    size_t tag;
                    if (!(in >> tag)) throw std::runtime_error("malformed element");
                    long long v[8] = {};
                    for (int q = 0; q < per; ++q)
                        if (!(in >> v[q])) throw std::runtime_error("malformed element");
                    if (type == 4) {
                        Int4 t;
                        for (int q = 0; q < 4; ++q) {
                            if (v[q] < 1 || v[q] > static_cast<long long>(vertices.size()))
                                throw std::runtime_error("MSH element node out of range");
                            t[q] = static_cast<int>(v[q] - 1);
                        }
                        tets.push_back(t);
                    }
                }
            }
        }
    }
    if (!formatOk) throw std::runtime_error("not an MSH file: " + filePath);
    return {std::move(vertices), std::move(tets)};
}

}  // namespace spg::io
