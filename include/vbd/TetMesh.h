// Source-compatible TetMesh (reference TetMesh.h:19-349) on the Eigen-free types.
//
// Same fields and methods: per-vertex state (3 x N column-major), tets, rest volumes, per-tet
// mu/lambda, vertex->tet adjacency, pinned flags; initialize (mass = rho V / 4 per corner,
// adjacency in ascending tet order, computed by the library's vbd_mesh_initialize), the
// time-step helpers (updateVelocity with the 0.999 damping, applyGravity, computeInertia,
// savePreviousPosition), and the boundary-surface OBJ export (TetMesh.h:45-186).
#pragma once

#include <algorithm>
#include <cstdint>
#include <fstream>
#include <string>
#include <vector>

#include "../vbd_capi.h"
#include "Types.h"

// The per-vertex time-step helpers run their loops on OpenMP threads when the including
// translation unit is compiled with OpenMP (each element's arithmetic is the same either way,
// so the results are bitwise those of one thread); without it they are plain loops.
#if defined(_OPENMP)
#define VBD_PER_VERTEX_FOR _Pragma("omp parallel for schedule(static) if (numVerts >= 65536)")
#else
#define VBD_PER_VERTEX_FOR
#endif

namespace VBD {

struct TetMesh {
    TVerticesMat mRestPos;
    TVerticesMat mVertPos;
    TVerticesMat mVertPrevPos;
    TVerticesMat mVelocity;
    TVerticesMat mInertia;
    VecDynamic vertexMass;

    std::vector<TetVIdsArr> tets;
    std::vector<FloatingType> tetRestVol;
    std::vector<FloatingType> tetMu;
    std::vector<FloatingType> tetLambda;

    std::vector<std::vector<IdType>> vertAdjacentTets;

    size_t numVerts = 0;
    size_t numTets = 0;

    std::vector<bool> isFixed;

    Vec3 vertex(IdType i) const { return mVertPos.col(i); }
    TVerticesMat::Col vertex(IdType i) { return mVertPos.col(i); }

    // TetMesh::initialize (TetMesh.h:213-275)
    void initialize(const std::vector<Vec3>& rest_positions, const std::vector<Vec3>& current_positions,
                    FloatingType density, const std::vector<TetVIdsArr>& tetIndices, FloatingType mu,
                    FloatingType lambda, const std::vector<IdType>& fixedVerts = {}) {
        if (rest_positions.size() != current_positions.size())
            throw std::invalid_argument("rest and current positions differ in size");
        numVerts = rest_positions.size();
        numTets = tetIndices.size();
        mRestPos.resize(3, numVerts);
        mVertPos.resize(3, numVerts);
        mVertPrevPos.resize(3, numVerts);
        mVelocity.resize(3, numVerts);
        mInertia.resize(3, numVerts);
        vertexMass.resize(numVerts);
        for (size_t i = 0; i < numVerts; ++i) {
            mRestPos.col(i) = rest_positions[i];
            mVertPos.col(i) = current_positions[i];
            mVertPrevPos.col(i) = current_positions[i];
        }
        tets = tetIndices;
        tetRestVol.assign(numTets, 0.f);
        tetMu.assign(numTets, mu);
        tetLambda.assign(numTets, lambda);
        std::vector<int32_t> off(numVerts + 1), adj(4 * numTets + 1);
        const int rc = vbd_mesh_initialize((int32_t)numVerts, mRestPos.data(), (int32_t)numTets,
                                           numTets ? tets[0].data() : nullptr, density, vertexMass.data(),
                                           tetRestVol.data(), off.data(), adj.data());
        if (rc != VBD_OK) throw std::runtime_error(std::string("vbd_mesh_initialize: ") + vbd_global_last_error());
        vertAdjacentTets.assign(numVerts, {});
        for (size_t v = 0; v < numVerts; ++v) vertAdjacentTets[v].assign(adj.begin() + off[v], adj.begin() + off[v + 1]);
        isFixed.assign(numVerts, false);
        for (IdType vid : fixedVerts)
            if (vid >= 0 && vid < (IdType)numVerts) isFixed[vid] = true;
    }

    FloatingType computeTetVolume(IdType tetId) const {
        const auto& v = tets[tetId];
        const Vec3 e0 = mVertPos.col(v[1]) - mVertPos.col(v[0]);
        const Vec3 e1 = mVertPos.col(v[2]) - mVertPos.col(v[0]);
        const Vec3 e2 = mVertPos.col(v[3]) - mVertPos.col(v[0]);
        return std::abs(e0.dot(e1.cross(e2))) / 6.0f;
    }

    FloatingType computeRestVolume(IdType tetId) const {
        const auto& v = tets[tetId];
        const Vec3 e0 = mRestPos.col(v[1]) - mRestPos.col(v[0]);
        const Vec3 e1 = mRestPos.col(v[2]) - mRestPos.col(v[0]);
        const Vec3 e2 = mRestPos.col(v[3]) - mRestPos.col(v[0]);
        return std::abs(e0.dot(e1.cross(e2))) / 6.0f;
    }

    // TetMesh.h:305-311
    void updateVelocity(FloatingType dt) {
        VBD_PER_VERTEX_FOR
        for (size_t i = 0; i < numVerts; ++i)
            if (!isFixed[i]) {
                auto v = mVelocity.col(i);
                for (int c = 0; c < 3; ++c) v[c] = ((mVertPos(c, i) - mVertPrevPos(c, i)) / dt) * 0.999f;
            }
    }

    // TetMesh.h:319-325
    void applyGravity(const Vec3& gravity, FloatingType dt) {
        const Vec3 dv = dt * gravity;
        VBD_PER_VERTEX_FOR
        for (size_t i = 0; i < numVerts; ++i)
            if (!isFixed[i]) mVelocity.col(i) += dv;
    }

    // TetMesh.h:332-341
    void computeInertia(FloatingType dt) {
        VBD_PER_VERTEX_FOR
        for (size_t i = 0; i < numVerts; ++i) {
            auto y = mInertia.col(i);
            for (int c = 0; c < 3; ++c) y[c] = isFixed[i] ? mVertPos(c, i) : mVertPos(c, i) + dt * mVelocity(c, i);
        }
    }

    void savePreviousPosition() {
        if (mVertPrevPos.cols() != mVertPos.cols()) {
            mVertPrevPos = mVertPos;
            return;
        }
        FloatingType* d = mVertPrevPos.data();
        const FloatingType* x = mVertPos.data();
        VBD_PER_VERTEX_FOR
        for (size_t i = 0; i < numVerts; ++i)
            for (int c = 0; c < 3; ++c) d[3 * i + c] = x[3 * i + c];
    }

    // Boundary faces: faces of the tets that occur exactly once, in ascending sorted-vertex
    // order, each with the winding of its (only) occurrence (TetMesh.h:80-141).
    std::vector<Vec3I> extractSurfaceFaces() const {
        // all 4T faces keyed by their sorted vertex triple, sorted once (instead of a tree of
        // 4T insertions); a key that occurs once is a boundary face
        struct Rec {
            IdType k[3];
            uint32_t face;  // 4 * tet + local face
        };
        std::vector<Rec> recs;
        recs.reserve(4 * tets.size());
        auto tri = [&](size_t t, int j) {
            const TetVIdsArr& v = tets[t];
            switch (j) {
                case 0: return Vec3I(v[0], v[2], v[1]);
                case 1: return Vec3I(v[0], v[1], v[3]);
                case 2: return Vec3I(v[0], v[3], v[2]);
                default: return Vec3I(v[1], v[2], v[3]);
            }
        };
        for (size_t t = 0; t < tets.size(); ++t)
            for (int j = 0; j < 4; ++j) {
                const Vec3I f = tri(t, j);
                Rec r{{f[0], f[1], f[2]}, (uint32_t)(4 * t + j)};
                std::sort(r.k, r.k + 3);
                recs.push_back(r);
            }
        std::sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) {
            if (a.k[0] != b.k[0]) return a.k[0] < b.k[0];
            if (a.k[1] != b.k[1]) return a.k[1] < b.k[1];
            if (a.k[2] != b.k[2]) return a.k[2] < b.k[2];
            return a.face < b.face;
        });
        std::vector<Vec3I> out;
        for (size_t i = 0; i < recs.size();) {
            size_t j = i + 1;
            while (j < recs.size() && recs[j].k[0] == recs[i].k[0] && recs[j].k[1] == recs[i].k[1] &&
                   recs[j].k[2] == recs[i].k[2])
                ++j;
            if (j == i + 1) out.push_back(tri(recs[i].face / 4, (int)(recs[i].face % 4)));
            i = j;
        }
        return out;
    }

    bool exportToObj(const std::string& filepath) const {
        std::ofstream f(filepath.c_str());
        if (!f.is_open()) return false;
        f << "# TetMesh OBJ Export\n# Vertices: " << numVerts << "\n# Tetrahedra: " << numTets << "\n\n";
        for (size_t i = 0; i < numVerts; ++i)
            f << "v " << mVertPos(0, i) << " " << mVertPos(1, i) << " " << mVertPos(2, i) << "\n";
        const auto faces = extractSurfaceFaces();
        f << "\n# Surface Triangles: " << faces.size() << "\n";
        for (const auto& t : faces) f << "f " << (t[0] + 1) << " " << (t[1] + 1) << " " << (t[2] + 1) << "\n";
        return true;
    }
};

}  // namespace VBD
