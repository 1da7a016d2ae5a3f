// Source-compatible TetMesh (reference TetMesh.h:19-349) on the Eigen-free types.
//
// Same fields and methods: per-vertex state (3 x N column-major), tets, rest volumes, per-tet
// mu/lambda, vertex->tet adjacency, pinned flags; initialize (mass = rho V / 4 per corner,
// adjacency in ascending tet order, computed by the library's vbd_mesh_initialize), the
// time-step helpers (updateVelocity with the 0.999 damping, applyGravity, computeInertia,
// savePreviousPosition), and the boundary-surface OBJ export (TetMesh.h:45-186).
#pragma once

#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../vbd_capi.h"
#include "Types.h"

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
        for (size_t i = 0; i < numVerts; ++i)
            if (!isFixed[i]) {
                auto v = mVelocity.col(i);
                for (int c = 0; c < 3; ++c) v[c] = ((mVertPos(c, i) - mVertPrevPos(c, i)) / dt) * 0.999f;
            }
    }

    // TetMesh.h:319-325
    void applyGravity(const Vec3& gravity, FloatingType dt) {
        for (size_t i = 0; i < numVerts; ++i)
            if (!isFixed[i]) mVelocity.col(i) += dt * gravity;
    }

    // TetMesh.h:332-341
    void computeInertia(FloatingType dt) {
        for (size_t i = 0; i < numVerts; ++i) {
            auto y = mInertia.col(i);
            for (int c = 0; c < 3; ++c) y[c] = isFixed[i] ? mVertPos(c, i) : mVertPos(c, i) + dt * mVelocity(c, i);
        }
    }

    void savePreviousPosition() { mVertPrevPos = mVertPos; }

    // Boundary faces: faces of the tets that occur exactly once, in ascending sorted-vertex
    // order, each with the winding of its (only) occurrence (TetMesh.h:80-141).
    std::vector<Vec3I> extractSurfaceFaces() const {
        struct Rec {
            int count;
            Vec3I oriented;
        };
        std::map<std::array<IdType, 3>, Rec> faces;
        for (const auto& t : tets) {
            const Vec3I tri[4] = {Vec3I(t[0], t[2], t[1]), Vec3I(t[0], t[1], t[3]), Vec3I(t[0], t[3], t[2]),
                                  Vec3I(t[1], t[2], t[3])};
            for (const Vec3I& f : tri) {
                std::array<IdType, 3> key = {f[0], f[1], f[2]};
                std::sort(key.begin(), key.end());
                auto it = faces.find(key);
                if (it == faces.end()) faces.emplace(key, Rec{1, f});
                else it->second.count++;
            }
        }
        std::vector<Vec3I> out;
        for (const auto& kv : faces)
            if (kv.second.count == 1) out.push_back(kv.second.oriented);
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
