// The drop-in TetMesh's per-vertex time-step helpers (TetMesh.h:305-348: applyGravity,
// computeInertia, updateVelocity, savePreviousPosition) on the arrays of an input file, their
// results written out for comparison with the float32 reference forms; built with and without
// OpenMP by the test.
//
// usage: host_steps_check <in.bin> <out.bin>
//   in:  int32 n, float x[3n], x_prev[3n], v[3n], uint8 fixed[n], float dt, float g[3]
//   out: float v after applyGravity[3n], y[3n], v after updateVelocity[3n], x_prev after
//        savePreviousPosition[3n]
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "vbd/TetMesh.h"

using namespace VBD;

static void put(TVerticesMat& m, const std::vector<float>& a, size_t n) {
    m.resize(3, n);
    for (size_t i = 0; i < 3 * n; ++i) m.data()[i] = a[i];
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    std::ifstream f(argv[1], std::ios::binary);
    int32_t n = 0;
    f.read((char*)&n, 4);
    std::vector<float> x(3 * (size_t)n), xp(3 * (size_t)n), v(3 * (size_t)n);
    std::vector<uint8_t> fixed(n);
    float dt = 0, g[3] = {0, 0, 0};
    f.read((char*)x.data(), 12 * (size_t)n);
    f.read((char*)xp.data(), 12 * (size_t)n);
    f.read((char*)v.data(), 12 * (size_t)n);
    f.read((char*)fixed.data(), n);
    f.read((char*)&dt, 4);
    f.read((char*)g, 12);
    if (!f) return 3;
    TetMesh mesh;
    mesh.numVerts = (size_t)n;
    put(mesh.mVertPos, x, n);
    put(mesh.mVertPrevPos, xp, n);
    put(mesh.mVelocity, v, n);
    put(mesh.mInertia, std::vector<float>(3 * (size_t)n, 0.f), n);
    mesh.isFixed.assign(n, false);
    for (int32_t i = 0; i < n; ++i) mesh.isFixed[i] = fixed[i] != 0;
    std::ofstream o(argv[2], std::ios::binary);
    mesh.applyGravity(Vec3(g[0], g[1], g[2]), dt);
    o.write((const char*)mesh.mVelocity.data(), 12 * (size_t)n);
    mesh.computeInertia(dt);
    o.write((const char*)mesh.mInertia.data(), 12 * (size_t)n);
    mesh.updateVelocity(dt);
    o.write((const char*)mesh.mVelocity.data(), 12 * (size_t)n);
    mesh.savePreviousPosition();
    o.write((const char*)mesh.mVertPrevPos.data(), 12 * (size_t)n);
    return o ? 0 : 4;
}
