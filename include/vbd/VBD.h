// Source-compatible VBDsolver / VBDParams (reference VBD.h:14-140), header-only over the C ABI.
//
// step(mesh) runs the whole VBD time step of VBD.cpp:60-101 on the GPU (forwardStep, the
// color sweeps with Chebyshev acceleration, updateVelocity) as one replayed hipGraph and
// leaves mesh.mVertPos / mVelocity / mInertia / mVertPrevPos updated like the reference.
// The other public members keep their reference meaning:
//   setParams, initializeColoring, precomputeTetData, simulate, computeTotalEnergy (device
//   reduction, |det F| for the singular-value product), getters, forwardStep (host),
//   solve (one device sweep on mesh.mVertPos towards mesh.mInertia), solveVertex and
//   computeTetContribution (host, closed-form block), applyAcceleration (host),
//   getAcceleratorOmega.
// Semantics follow VBD.cpp with numThreads = 1: colors in ascending order (the reference's
// OpenMP loop over colors at VBD.cpp:120-121 races; numThreads is accepted and ignored).
// Extensions (no counterpart): usePerTetMaterial, velocityDamping, contact*, and
// stepDevice()/syncToHost() for loops that keep the state on the device.
#pragma once

#include <functional>
#include <string>
#include <vector>

#include "../vbd_capi.h"
#include "GraphColoring.h"
#include "TetMesh.h"
#include "Types.h"
#include "VBD_GPU_Wrapper.h"

namespace VBD {

struct VBDParams {
    Real dt = 0.01666666f;
    int numIterations = 50;
    int substeps = 1;
    Vec3 gravity{0, -9.8f, 0};
    bool useAcceleration = false;
    Real accelerationRho = 0.75f;
    int numThreads = 4;
    Real mu = 1e6f;
    Real lambda = 1e6f;
    Real density = 1000.0f;
    // extensions
    bool usePerTetMaterial = false;  // use TetMesh::tetMu / tetLambda instead of mu / lambda
    Real velocityDamping = 0.999f;   // TetMesh.h:308
    bool contactOn = false;          // penalty half-space x[contactAxis] >= contactPlane
    int contactAxis = 2;
    Real contactStiffness = 1e4f;
    Real contactPlane = 0.f;
    int energyLogEvery = 10;  // step() records the diagnostic energy every n-th iteration (VBD.cpp:92-94); 0: off
};

class VBDsolver {
   public:
    VBDsolver() = default;
    ~VBDsolver() { vbd_destroy(h_); }
    VBDsolver(const VBDsolver&) = delete;
    VBDsolver& operator=(const VBDsolver&) = delete;

    void setParams(const VBDParams& p) {
        params = p;
        if (h_) applyParams();
    }

    void initializeColoring(const TetMesh& mesh) {
        vertexColors = GraphColoring::colorVertices(mesh.vertAdjacentTets, mesh.tets, mesh.numVerts);
        prevPrevPos.resize(3, mesh.numVerts);
        prevPrevPos.setZero();
    }

    void precomputeTetData(const TetMesh& mesh) {
        std::vector<float> d(9 * mesh.numTets + 9);
        tetRestVolume.assign(mesh.numTets, 0.f);
        const int rc = vbd_precompute_tet_data(mesh.mRestPos.data(), (int32_t)mesh.numTets,
                                               mesh.numTets ? mesh.tets[0].data() : nullptr, d.data(),
                                               tetRestVolume.data());
        if (rc != VBD_OK) throw std::runtime_error(std::string("vbd_precompute_tet_data: ") + vbd_global_last_error());
        tetDmInv.assign(mesh.numTets, Matrix3r());
        for (size_t t = 0; t < mesh.numTets; ++t)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) tetDmInv[t](a, b) = d[9 * t + 3 * a + b];
    }

    // VBD.cpp:60-101
    void step(TetMesh& mesh) {
        ensureDevice(mesh);
        vbd_throw_if(vbd_upload_positions(h_, mesh.mVertPos.data()), h_, "step");
        vbd_throw_if(vbd_upload_velocities(h_, mesh.mVelocity.data()), h_, "step");
        stepDevice(1);
        syncToHost(mesh);
        if (params.energyLogEvery > 0) {  // VBD.cpp:92-94: energies of this step, in order
            int32_t n = 0;
            vbd_throw_if(vbd_energy_log(h_, nullptr, 0, &n), h_, "energy log");
            std::vector<double> e((size_t)n);
            vbd_throw_if(vbd_energy_log(h_, e.data(), n, &n), h_, "energy log");
            for (double v : e) energyHistory.push_back((Real)v);
        }
    }

    // Advance `frames` steps without host transfers (state stays on the device).
    void stepDevice(int frames) {
        if (!h_) throw std::runtime_error("stepDevice before the first step(mesh)");
        vbd_throw_if(vbd_step(h_, frames), h_, "vbd_step");
    }

    // Download positions, velocities, the inertia target and x_prev of the last substep.
    void syncToHost(TetMesh& mesh) {
        vbd_throw_if(vbd_download_positions(h_, mesh.mVertPos.data()), h_, "download positions");
        vbd_throw_if(vbd_download_velocities(h_, mesh.mVelocity.data()), h_, "download velocities");
        vbd_throw_if(vbd_download_inertia(h_, mesh.mInertia.data()), h_, "download inertia");
        vbd_throw_if(vbd_download_previous_positions(h_, mesh.mVertPrevPos.data()), h_, "download x_prev");
    }

    // VBD.cpp:313-330
    void simulate(TetMesh& mesh, int numFrames, std::function<void(int, TetMesh&)> callback = nullptr) {
        energyHistory.clear();
        for (int f = 0; f < numFrames; ++f) {
            step(mesh);
            if (callback) callback(f, mesh);
        }
    }

    // VBD.cpp:258-288 (device reduction)
    Real computeTotalEnergy(const TetMesh& mesh) const {
        VBDsolver* self = const_cast<VBDsolver*>(this);
        self->ensureDevice(mesh);
        vbd_throw_if(vbd_upload_positions(h_, mesh.mVertPos.data()), h_, "computeTotalEnergy");
        double e = 0;
        vbd_throw_if(vbd_total_energy(h_, params.mu, params.lambda, &e), h_, "computeTotalEnergy");
        return (Real)e;
    }

    const std::vector<std::vector<IdType>>& getVertexColors() const { return vertexColors; }
    const std::vector<Matrix3r>& getTetDmInv() const { return tetDmInv; }
    const std::vector<Real>& getTetRestVol() const { return tetRestVolume; }

    // VBD.cpp:103-112 (host, as the reference main.cpp:220 uses it)
    void forwardStep(TetMesh& mesh) {
        mesh.savePreviousPosition();
        mesh.applyGravity(params.gravity, params.dt / params.substeps);
        mesh.computeInertia(params.dt / params.substeps);
    }

    // One Gauss-Seidel sweep (VBD.cpp:114-158) on the device, from mesh.mVertPos towards
    // mesh.mInertia; the result is written back to mesh.mVertPos.
    void solve(TetMesh& mesh) {
        ensureDevice(mesh);
        vbd_throw_if(vbd_upload_positions(h_, mesh.mVertPos.data()), h_, "solve");
        vbd_throw_if(vbd_upload_inertia(h_, mesh.mInertia.data()), h_, "solve");
        const Real dt = params.dt / params.substeps;
        const Real s = (Real)(1.0 / (double)(dt * dt));
        vbd_throw_if(vbd_solve_iteration(h_, s, params.mu, params.lambda), h_, "solve");
        vbd_throw_if(vbd_download_positions(h_, mesh.mVertPos.data()), h_, "solve");
    }

    // VBD.cpp:160-203 on the host: H = m/h^2 I + sum_t H_kk(t), f = m/h^2 (y - x) - sum_t g_k(t).
    void solveVertex(const TetMesh& mesh, IdType vertexId, Mat3& H, Vec3& f, Real dtSqrReciprocal) {
        const Real m = mesh.vertexMass(vertexId);
        const Real ms = m * dtSqrReciprocal;
        H(0, 0) += ms; H(1, 1) += ms; H(2, 2) += ms;
        const Vec3 x = mesh.mVertPos.col(vertexId), y = mesh.mInertia.col(vertexId);
        for (int c = 0; c < 3; ++c) f[c] = (m * (y[c] - x[c])) * dtSqrReciprocal;
        for (IdType t : mesh.vertAdjacentTets[vertexId]) {
            int k = -1;
            for (int i = 0; i < 4; ++i)
                if (mesh.tets[t][i] == vertexId) { k = i; break; }
            if (k < 0) continue;
            Mat3 Hl;
            Vec3 fl;
            computeTetContribution(mesh, t, vertexId, k, Hl, fl);
            H += Hl;
            f -= fl;
        }
    }

    // VBD.cpp:205-256 in closed form: g_k = V P q_k, H_kk = V [mu |q|^2 I + lambda (C q)(C q)^T]
    // with F = Ds DmInv, C = cof F, J = det F, P = mu (F - C) + lambda (J - 1) C.
    void computeTetContribution(const TetMesh& mesh, IdType tetId, IdType vertexId, int localVertexIdx,
                                Mat3& H_local, Vec3& f_local) {
        (void)vertexId;
        const auto& tet = mesh.tets[tetId];
        const Real mu = params.usePerTetMaterial ? mesh.tetMu[tetId] : params.mu;
        const Real lam = params.usePerTetMaterial ? mesh.tetLambda[tetId] : params.lambda;
        const Real V = tetRestVolume[tetId];
        const Matrix3r& D = tetDmInv[tetId];
        Mat3 Ds, F, C;
        for (int j = 0; j < 3; ++j)
            for (int c = 0; c < 3; ++c) Ds(c, j) = mesh.mVertPos(c, tet[j + 1]) - mesh.mVertPos(c, tet[0]);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) F(r, c) = Ds(r, 0) * D(0, c) + (Ds(r, 1) * D(1, c) + Ds(r, 2) * D(2, c));
        C(0, 0) = F(1, 1) * F(2, 2) - F(1, 2) * F(2, 1);
        C(0, 1) = F(1, 2) * F(2, 0) - F(1, 0) * F(2, 2);
        C(0, 2) = F(1, 0) * F(2, 1) - F(1, 1) * F(2, 0);
        C(1, 0) = F(2, 1) * F(0, 2) - F(2, 2) * F(0, 1);
        C(1, 1) = F(2, 2) * F(0, 0) - F(2, 0) * F(0, 2);
        C(1, 2) = F(2, 0) * F(0, 1) - F(2, 1) * F(0, 0);
        C(2, 0) = F(0, 1) * F(1, 2) - F(0, 2) * F(1, 1);
        C(2, 1) = F(0, 2) * F(1, 0) - F(0, 0) * F(1, 2);
        C(2, 2) = F(0, 0) * F(1, 1) - F(0, 1) * F(1, 0);
        const Real J = F.determinant();
        const Real a = (Real)((double)lam * ((double)J - 1.0));
        Real q[3];
        for (int b = 0; b < 3; ++b) q[b] = localVertexIdx == 0 ? -(D(0, b) + (D(1, b) + D(2, b))) : D(localVertexIdx - 1, b);
        Real Cq[3], Pq[3];
        for (int r = 0; r < 3; ++r) {
            Real cq = 0, pq = 0;
            for (int b = 0; b < 3; ++b) {
                cq += C(r, b) * q[b];
                pq += (mu * (F(r, b) - C(r, b)) + a * C(r, b)) * q[b];
            }
            Cq[r] = cq;
            Pq[r] = pq;
        }
        const Real qq = q[0] * q[0] + (q[1] * q[1] + q[2] * q[2]);
        for (int r = 0; r < 3; ++r) {
            f_local[r] = V * Pq[r];
            for (int c = 0; c < 3; ++c) H_local(r, c) = V * ((r == c ? mu * qq : Real(0)) + lam * Cq[r] * Cq[c]);
        }
    }

    // VBD.cpp:290-300 (host)
    void applyAcceleration(TetMesh& mesh, Real omega) {
        if (omega <= 1.0) return;
        for (size_t i = 0; i < mesh.numVerts; ++i)
            if (!mesh.isFixed[i])
                for (int c = 0; c < 3; ++c)
                    mesh.mVertPos(c, i) = omega * (mesh.mVertPos(c, i) - prevPrevPos(c, i)) + prevPrevPos(c, i);
    }

    // VBD.cpp:302-311
    Real getAcceleratorOmega(int order, Real rho, Real prevOmega) { return vbd_accelerator_omega(order, rho, prevOmega); }

    vbd_solver* handle() const { return h_; }

   public:
    VBDParams params;
    std::vector<std::vector<IdType>> vertexColors;
    TVerticesMat prevPrevPos;
    std::vector<Real> energyHistory;
    std::vector<Matrix3r> tetDmInv;
    std::vector<Real> tetRestVolume;

   private:
    void applyParams() {
        vbd_step_params p{};
        p.dt = params.dt;
        p.num_iterations = params.numIterations;
        p.substeps = params.substeps;
        for (int c = 0; c < 3; ++c) p.gravity[c] = params.gravity[c];
        p.use_acceleration = params.useAcceleration ? 1 : 0;
        p.acceleration_rho = params.accelerationRho;
        p.mu = params.mu;
        p.lambda = params.lambda;
        p.velocity_damping = params.velocityDamping;
        p.start_from_inertia = 0;
        p.contact_on = params.contactOn ? 1 : 0;
        p.contact_axis = params.contactAxis;
        p.contact_k = params.contactStiffness;
        p.contact_plane = params.contactPlane;
        vbd_throw_if(vbd_set_params(h_, &p), h_, "setParams");
        vbd_throw_if(vbd_set_energy_log(h_, params.energyLogEvery), h_, "setParams");
    }

    void ensureDevice(const TetMesh& mesh) {
        if (vertexColors.empty()) {
            initializeColoring(mesh);
            precomputeTetData(mesh);
        }
        if (h_) return;
        vbd_throw_if(vbd_create(&h_, 0), nullptr, "vbd_create");
        std::vector<uint8_t> fixed(mesh.numVerts);
        for (size_t i = 0; i < mesh.numVerts; ++i) fixed[i] = mesh.isFixed[i] ? 1 : 0;
        std::vector<float> dminv(9 * mesh.numTets);
        for (size_t t = 0; t < mesh.numTets; ++t)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) dminv[9 * t + 3 * a + b] = tetDmInv[t](a, b);
        std::vector<int32_t> coff(vertexColors.size() + 1, 0), cverts;
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
        d.rest_volume = tetRestVolume.data();
        d.tet_mu = params.usePerTetMaterial ? mesh.tetMu.data() : nullptr;
        d.tet_lambda = params.usePerTetMaterial ? mesh.tetLambda.data() : nullptr;
        d.num_colors = (int32_t)vertexColors.size();
        d.color_offsets = coff.data();
        d.color_verts = cverts.data();
        vbd_throw_if(vbd_init(h_, &d), h_, "upload");
        applyParams();
    }

    vbd_solver* h_ = nullptr;
};

}  // namespace VBD
