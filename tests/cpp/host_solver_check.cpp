// Runs the host-side VBDsolver helpers of include/vbd/VBD.h (the reference VBD.cpp:36-58,
// :103-112 and :160-256 on the host) on a TetGen model and writes what they produced, for
// tests/test_dropin.py:
//   int32 nV, int32 nC, int32 color_of[nV] (initializeColoring),
//   float y[3 nV] (forwardStep from rest velocity), float H[9 nV] and f[3 nV] (solveVertex at
//   the current positions towards y; zero for pinned vertices).
// usage: host_solver_check <prefix> <current node file> <mu> <lambda> <out>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "vbd/TetGenIO.h"
#include "vbd/VBD.h"

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: host_solver_check <prefix> <current node file> <mu> <lambda> <out>\n");
        return 2;
    }
    const std::string prefix = argv[1];
    const Real mu = (Real)std::atof(argv[3]), lambda = (Real)std::atof(argv[4]);
    VBD::TetMesh mesh;
    if (!VBD::loadTetgenModel(prefix + ".node", argv[2], prefix + ".ele", mesh, 1000.0f, mu, lambda, 0, 0.01f)) {
        std::fprintf(stderr, "load failed\n");
        return 1;
    }
    VBD::VBDParams p;
    p.dt = 1.0f / 60.0f;
    p.substeps = 1;
    p.mu = mu;
    p.lambda = lambda;
    VBD::VBDsolver s;
    s.setParams(p);
    s.initializeColoring(mesh);
    s.precomputeTetData(mesh);
    s.forwardStep(mesh);
    const Real h = p.dt / p.substeps;
    const Real sdt = (Real)(1.0 / (double)(h * h));
    const int32_t nV = (int32_t)mesh.numVerts;
    std::vector<int32_t> colorOf(nV, -1);
    const auto& colors = s.getVertexColors();
    for (size_t c = 0; c < colors.size(); ++c)
        for (VBD::IdType v : colors[c]) colorOf[v] = (int32_t)c;
    std::vector<float> H(9 * (size_t)nV, 0.f), f(3 * (size_t)nV, 0.f);
    for (int32_t v = 0; v < nV; ++v) {
        if (mesh.isFixed[v]) continue;
        VBD::Mat3 Hv;
        VBD::Vec3 fv;
        s.solveVertex(mesh, v, Hv, fv, sdt);
        for (int r = 0; r < 3; ++r) {
            f[3 * (size_t)v + r] = fv[r];
            for (int c = 0; c < 3; ++c) H[9 * (size_t)v + 3 * r + c] = Hv(r, c);
        }
    }
    FILE* out = std::fopen(argv[5], "wb");
    if (!out) return 3;
    const int32_t nC = (int32_t)colors.size();
    std::fwrite(&nV, 4, 1, out);
    std::fwrite(&nC, 4, 1, out);
    std::fwrite(colorOf.data(), 4, (size_t)nV, out);
    std::fwrite(mesh.mInertia.data(), sizeof(float), 3 * (size_t)nV, out);
    std::fwrite(H.data(), sizeof(float), H.size(), out);
    std::fwrite(f.data(), sizeof(float), f.size(), out);
    std::fclose(out);
    return 0;
}
