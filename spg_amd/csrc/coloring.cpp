#include "coloring.h"

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

}  // namespace spgvbd
