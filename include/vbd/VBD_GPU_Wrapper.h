// Source-compatible VBD_GPU_Solver (reference VBD_GPU_Wrapper.h:7-45), header-only over the
// C ABI in vbd_capi.h.  Same public methods and argument meaning:
//   initFromCPU(mesh, vertexColors, tetDmInv, tetRestVol)   VBD_GPU_Wrapper.cu:246-315
//   solveIteration(dtSqrReciprocal, mu, lambda)             VBD_GPU_Wrapper.cu:317-329
//   solveVBD(numIterations, dtSqrReciprocal, mu, lambda, useAcceleration = true, rho = 0.99f)
//                                                           VBD_GPU_Wrapper.cu:332-362
//   copyPositionsToCPU(mesh)                                VBD_GPU_Wrapper.cu:364-366
//   updateInertiaFromCPU(mesh)                              VBD_GPU_Wrapper.cu:369-371
// Differences, by design: the default element model is the CPU solver's stable Neo-Hookean
// (Energy.cpp:90-158) without the 1.000001 inertia regularisation; setMaterialModel(
// VBD_MATERIAL_COROTATED_GPU) selects the corotated model the reference kernel calls, with
// that regularisation.  Failures throw std::runtime_error instead of being ignored; the
// object owns its device memory and is movable but not copyable.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "../vbd_capi.h"
#include "TetMesh.h"

namespace VBD {

inline void vbd_throw_if(int rc, const vbd_solver* h, const char* what) {
    if (rc != VBD_OK)
        throw std::runtime_error(std::string(what) + ": " + (h ? vbd_last_error(h) : vbd_global_last_error()));
}

class VBD_GPU_Solver {
   public:
    explicit VBD_GPU_Solver(int device = 0) { vbd_throw_if(vbd_create(&h_, device), nullptr, "vbd_create"); }
    ~VBD_GPU_Solver() { vbd_destroy(h_); }
    VBD_GPU_Solver(const VBD_GPU_Solver&) = delete;
    VBD_GPU_Solver& operator=(const VBD_GPU_Solver&) = delete;
    VBD_GPU_Solver(VBD_GPU_Solver&& o) noexcept : h_(std::exchange(o.h_, nullptr)), numVerts_(o.numVerts_) {}
    VBD_GPU_Solver& operator=(VBD_GPU_Solver&& o) noexcept {
        if (this != &o) {
            vbd_destroy(h_);
            h_ = std::exchange(o.h_, nullptr);
            numVerts_ = o.numVerts_;
        }
        return *this;
    }

    // Extension: VBD_MATERIAL_COROTATED_GPU reproduces the reference GPU kernel's corotated model
    // (VBD_GPU_Wrapper.cu:181); the default is the CPU solver's stable Neo-Hookean.
    void setMaterialModel(int model) { vbd_throw_if(vbd_set_material_model(h_, model), h_, "setMaterialModel"); }

    void initFromCPU(const TetMesh& mesh, const std::vector<std::vector<int>>& vertexColors,
                     const std::vector<Matrix3r>& tetDmInv, const std::vector<Real>& tetRestVol) {
        std::vector<uint8_t> fixed(mesh.numVerts);
        for (size_t i = 0; i < mesh.numVerts; ++i) fixed[i] = mesh.isFixed[i] ? 1 : 0;
        std::vector<float> dminv(9 * mesh.numTets);
        for (size_t t = 0; t < mesh.numTets; ++t)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) dminv[9 * t + 3 * a + b] = tetDmInv[t](a, b);
        std::vector<int32_t> coff(vertexColors.size() + 1, 0), cverts;
        cverts.reserve(mesh.numVerts);
        for (size_t c = 0; c < vertexColors.size(); ++c) {
            cverts.insert(cverts.end(), vertexColors[c].begin(), vertexColors[c].end());
            coff[c + 1] = (int32_t)cverts.size();
        }
        vbd_mesh_desc d{};
        d.num_verts = (int32_t)mesh.numVerts;
        d.num_tets = (int32_t)mesh.numTets;
        d.positions = mesh.mVertPos.data();
        d.inertia = mesh.mInertia.data();
        d.velocity = mesh.mVelocity.data();
        d.mass = mesh.vertexMass.data();
        d.is_fixed = fixed.data();
        d.tets = mesh.numTets ? mesh.tets[0].data() : nullptr;
        d.dminv = dminv.data();
        d.rest_volume = tetRestVol.data();
        d.num_colors = (int32_t)vertexColors.size();
        d.color_offsets = coff.data();
        d.color_verts = cverts.data();
        vbd_throw_if(vbd_init(h_, &d), h_, "initFromCPU");
        numVerts_ = mesh.numVerts;
    }

    void solveIteration(Real dtSqrReciprocal, Real mu, Real lambda) {
        vbd_throw_if(vbd_solve_iteration(h_, dtSqrReciprocal, mu, lambda), h_, "solveIteration");
    }

    void copyPositionsToCPU(TetMesh& mesh) {
        vbd_throw_if(vbd_download_positions(h_, mesh.mVertPos.data()), h_, "copyPositionsToCPU");
    }

    void updateInertiaFromCPU(const TetMesh& mesh) {
        vbd_throw_if(vbd_upload_inertia(h_, mesh.mInertia.data()), h_, "updateInertiaFromCPU");
    }

    // x <- y, then numIterations sweeps with Chebyshev acceleration (VBD_GPU_Wrapper.cu:332-362)
    void solveVBD(int numIterations, Real dtSqrReciprocal, Real mu, Real lambda, bool useAcceleration = true,
                  Real rho = 0.99f) {
        vbd_throw_if(vbd_solve(h_, numIterations, dtSqrReciprocal, mu, lambda, useAcceleration ? 1 : 0, rho, 1), h_,
                     "solveVBD");
    }

    vbd_solver* handle() const { return h_; }

   private:
    vbd_solver* h_ = nullptr;
    size_t numVerts_ = 0;
};

}  // namespace VBD
