// Source-compatible GraphColoring (reference GraphColoring.h:12-44): greedy first-fit
// vertex coloring in ascending vertex order over the graph of shared tets
// (GraphColoring.cpp:28-66), computed by the library's host helper vbd_color_vertices.
#pragma once

#include <string>
#include <vector>

#include "../vbd_capi.h"
#include "Types.h"

namespace VBD {

class GraphColoring {
   public:
    static std::vector<std::vector<IdType>> colorVertices(const std::vector<std::vector<IdType>>& vertAdjacentTets,
                                                          const std::vector<TetVIdsArr>& tets, size_t numVerts) {
        std::vector<int32_t> off(numVerts + 1, 0), adj;
        for (size_t v = 0; v < numVerts; ++v) {
            adj.insert(adj.end(), vertAdjacentTets[v].begin(), vertAdjacentTets[v].end());
            off[v + 1] = (int32_t)adj.size();
        }
        std::vector<int32_t> colorOf(numVerts + 1), coff(numVerts + 2), cverts(numVerts + 1);
        int32_t nc = 0;
        const int rc = vbd_color_vertices((int32_t)numVerts, tets.empty() ? nullptr : tets[0].data(), off.data(),
                                          adj.empty() ? off.data() : adj.data(), colorOf.data(), coff.data(),
                                          cverts.data(), &nc);
        if (rc != VBD_OK) throw std::runtime_error(std::string("vbd_color_vertices: ") + vbd_global_last_error());
        std::vector<std::vector<IdType>> groups(nc);
        for (int c = 0; c < nc; ++c) groups[c].assign(cverts.begin() + coff[c], cverts.begin() + coff[c + 1]);
        return groups;
    }

    static size_t getNumColors(const std::vector<std::vector<IdType>>& colors) { return colors.size(); }
};

}  // namespace VBD
