// Loads a TetGen model with include/vbd/TetGenIO.h (the loader of reference main.cpp:14-168)
// and writes what the loader produced to a binary file for tests/test_dropin.py:
//   int32 nV, int32 nT, float rest[3 nV], float x[3 nV], int32 tets[4 nT], uint8 fixed[nV],
//   float mass[nV], then the surface triangles: int32 nF, int32 faces[3 nF]; with a sixth
//   argument, also TetMesh::exportToObj to that path.
// usage: tetgen_io_check <prefix> <current-node file or "-"> <pin axis> <pin fraction> <out> [obj]
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "vbd/TetGenIO.h"

int main(int argc, char** argv) {
    if (argc != 6 && argc != 7) {
        std::fprintf(stderr, "usage: tetgen_io_check <prefix> <current|-> <pin axis> <pin fraction> <out>\n");
        return 2;
    }
    const std::string prefix = argv[1], cur = argv[2];
    VBD::TetMesh mesh;
    if (!VBD::loadTetgenModel(prefix + ".node", cur == "-" ? std::string() : cur, prefix + ".ele", mesh, 1000.0f,
                              1e5f, 4e5f, std::atoi(argv[3]), (Real)std::atof(argv[4]))) {
        std::fprintf(stderr, "load failed\n");
        return 1;
    }
    FILE* f = std::fopen(argv[5], "wb");
    if (!f) return 3;
    const int32_t nV = (int32_t)mesh.numVerts, nT = (int32_t)mesh.numTets;
    std::fwrite(&nV, 4, 1, f);
    std::fwrite(&nT, 4, 1, f);
    std::fwrite(mesh.mRestPos.data(), sizeof(float), 3 * (size_t)nV, f);
    std::fwrite(mesh.mVertPos.data(), sizeof(float), 3 * (size_t)nV, f);
    for (const auto& t : mesh.tets) std::fwrite(t.data(), sizeof(int32_t), 4, f);
    for (int32_t i = 0; i < nV; ++i) {
        const uint8_t b = mesh.isFixed[i] ? 1 : 0;
        std::fwrite(&b, 1, 1, f);
    }
    std::fwrite(mesh.vertexMass.data(), sizeof(float), (size_t)nV, f);
    const auto faces = mesh.extractSurfaceFaces();
    const int32_t nF = (int32_t)faces.size();
    std::fwrite(&nF, 4, 1, f);
    for (const auto& fc : faces) {
        const int32_t v[3] = {fc[0], fc[1], fc[2]};
        std::fwrite(v, sizeof(int32_t), 3, f);
    }
    std::fclose(f);
    if (argc == 7 && !mesh.exportToObj(argv[6])) return 4;
    return 0;
}
