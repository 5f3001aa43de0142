#include "coloring.h"

#include <algorithm>
#include <stdexcept>

namespace spgvbd {

std::vector<int32_t> colorVertices(int64_t nv, const std::vector<FlatStencilSet> &sets)
{
    std::vector<int32_t> colors(static_cast<size_t>(nv > 0 ? nv : 0), -1);
    if (nv <= 0) return colors;
    // degree count, then fill: for every stencil, every unordered pair (i, j) with i before j
    std::vector<int64_t> deg(nv + 1, 0);
    for (const auto &s : sets) {
        if (s.stencilSize <= 0) throw std::runtime_error("invalid stencil size");
        for (int64_t i = 0; i < s.length; ++i) {
            const int32_t a = s.entries[i];
            if (a < 0 || a >= nv) throw std::runtime_error("Trying to add edge to graph with invalid node indices");
            for (int64_t j = i + 1; j % s.stencilSize != 0; ++j) {
                const int32_t b = s.entries[j];
                if (b < 0 || b >= nv)
                    throw std::runtime_error("Trying to add edge to graph with invalid node indices");
                ++deg[a + 1];
                ++deg[b + 1];
            }
        }
    }
    for (int64_t v = 0; v < nv; ++v) deg[v + 1] += deg[v];
    std::vector<int32_t> adj(static_cast<size_t>(deg[nv]));
    std::vector<int64_t> fill(deg.begin(), deg.end() - 1);
    for (const auto &s : sets)
        for (int64_t i = 0; i < s.length; ++i)
            for (int64_t j = i + 1; j % s.stencilSize != 0; ++j) {
                const int32_t a = s.entries[i], b = s.entries[j];
                adj[fill[a]++] = b;
                adj[fill[b]++] = a;
            }
    std::vector<char> used(static_cast<size_t>(nv) + 1, 0);
    colors[0] = 0;
    for (int64_t n = 1; n < nv; ++n) {
        for (int64_t k = deg[n]; k < deg[n + 1]; ++k) {
            const int32_t c = colors[adj[k]];
            if (c != -1) used[c] = 1;
        }
        int32_t c = 0;
        while (used[c]) ++c;
        colors[n] = c;
        for (int64_t k = deg[n]; k < deg[n + 1]; ++k) {
            const int32_t q = colors[adj[k]];
            if (q != -1) used[q] = 0;
        }
    }
    return colors;
}

std::vector<int32_t> colorStencils(const int32_t *flat, int64_t nStencils, int32_t stencilSize)
{
    std::vector<int32_t> colors(static_cast<size_t>(nStencils > 0 ? nStencils : 0), -1);
    if (nStencils <= 0) return colors;
    if (stencilSize <= 0) throw std::runtime_error("invalid stencil size");
    const int64_t len = nStencils * stencilSize;
    int32_t maxV = -1;
    for (int64_t k = 0; k < len; ++k) {
        if (flat[k] < 0) throw std::runtime_error("Trying to add edge to graph with invalid node indices");
        maxV = std::max(maxV, flat[k]);
    }
    // per-vertex stencil lists (CSR), each ascending in stencil index
    std::vector<int64_t> ptr(static_cast<size_t>(maxV) + 2, 0);
    for (int64_t k = 0; k < len; ++k) ++ptr[flat[k] + 1];
    for (int32_t v = 0; v <= maxV; ++v) ptr[v + 1] += ptr[v];
    std::vector<int32_t> list(static_cast<size_t>(len));
    std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t k = 0; k < len; ++k) list[fill[flat[k]]++] = static_cast<int32_t>(k / stencilSize);
    std::vector<char> used(static_cast<size_t>(nStencils) + 1, 0);
    colors[0] = 0;
    for (int64_t n = 1; n < nStencils; ++n) {
        const int32_t *st = flat + n * stencilSize;
        // the earlier stencils of each list are a prefix of it (lists ascend)
        for (int32_t j = 0; j < stencilSize; ++j)
            for (int64_t k = ptr[st[j]]; k < ptr[st[j] + 1] && list[k] < n; ++k) used[colors[list[k]]] = 1;
        int32_t c = 0;
        while (used[c]) ++c;
        colors[n] = c;
        for (int32_t j = 0; j < stencilSize; ++j)
            for (int64_t k = ptr[st[j]]; k < ptr[st[j] + 1] && list[k] < n; ++k) used[colors[list[k]]] = 0;
    }
    return colors;
}

std::vector<int32_t> serialLevels(const int32_t *flat, int64_t nStencils, int32_t stencilSize)
{
    std::vector<int32_t> level(static_cast<size_t>(nStencils > 0 ? nStencils : 0), 0);
    if (nStencils <= 0) return level;
    if (stencilSize <= 0) throw std::runtime_error("invalid stencil size");
    int32_t maxV = -1;
    for (int64_t k = 0; k < nStencils * stencilSize; ++k) {
        if (flat[k] < 0) throw std::runtime_error("negative particle index");
        maxV = std::max(maxV, flat[k]);
    }
    std::vector<int32_t> last(static_cast<size_t>(maxV) + 1, -1);  // level of the last stencil at each vertex
    for (int64_t s = 0; s < nStencils; ++s) {
        const int32_t *st = flat + s * stencilSize;
        int32_t l = 0;
        for (int32_t j = 0; j < stencilSize; ++j) l = std::max(l, last[st[j]] + 1);
        level[s] = l;
        for (int32_t j = 0; j < stencilSize; ++j) last[st[j]] = l;
    }
    return level;
}

}  // namespace spgvbd
