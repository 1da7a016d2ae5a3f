// Reference-style driver over the source-compatible C++ drop-in (include/vbd/*.h).
//
// Loads a TetGen model, prepares it exactly as reference main.cpp:188-208 does
// (VBDsolver::initializeColoring, precomputeTetData, VBD_GPU_Solver::initFromCPU), then runs
// either the reference's hybrid frame loop (main.cpp:215-232: host forwardStep,
// updateInertiaFromCPU, solveVBD, copyPositionsToCPU, host updateVelocity) or
// VBDsolver::step, and writes the positions after every frame as raw float32.
//
// usage: vbd_dropin_demo <model prefix> <hybrid|step> <frames> <iterations> <rho>
//                        <pin axis> <pin fraction> <mu> <lambda> <out.bin> [obj]
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "vbd/TetGenIO.h"
#include "vbd/VBD.h"
#include "vbd/VBD_GPU_Wrapper.h"

using namespace VBD;

int main(int argc, char** argv) {
    if (argc < 11) {
        std::cerr << "usage: vbd_dropin_demo <prefix> <hybrid|step> <frames> <iterations> <rho> <pin axis> "
                     "<pin fraction> <mu> <lambda> <out.bin> [obj]\n";
        return 2;
    }
    const std::string prefix = argv[1], mode = argv[2], out = argv[10];
    const int frames = std::atoi(argv[3]), iters = std::atoi(argv[4]);
    const Real rho = (Real)std::atof(argv[5]);
    const int pinAxis = std::atoi(argv[6]);
    const Real pinFrac = (Real)std::atof(argv[7]);
    const Real mu = (Real)std::atof(argv[8]), lambda = (Real)std::atof(argv[9]);
    try {
        TetMesh mesh;
        if (!loadTetgenModel(prefix + ".node", "", prefix + ".ele", mesh, 1000.0f, mu, lambda, pinAxis, pinFrac)) {
            std::cerr << "failed to load " << prefix << "\n";
            return 1;
        }
        VBDParams params;
        params.dt = 1.0f / 60.0f;
        params.numIterations = iters;
        params.substeps = 1;
        params.gravity = Vec3(0.0f, -9.8f, 0.0f);
        params.useAcceleration = true;
        params.accelerationRho = rho;
        params.mu = mu;
        params.lambda = lambda;
        std::ofstream f(out, std::ios::binary);
        VBDsolver cpuSolver;
        cpuSolver.setParams(params);
        cpuSolver.initializeColoring(mesh);
        cpuSolver.precomputeTetData(mesh);
        std::cout << "mesh " << mesh.numVerts << " verts, " << mesh.numTets << " tets, "
                  << cpuSolver.getVertexColors().size() << " colors" << std::endl;
        if (mode == "hybrid") {
            VBD_GPU_Solver gpuSolver;
            gpuSolver.initFromCPU(mesh, cpuSolver.getVertexColors(), cpuSolver.getTetDmInv(), cpuSolver.getTetRestVol());
            const Real subDt = params.dt / params.substeps;
            const Real dtSqrReciprocal = 1.0f / (subDt * subDt);
            for (int frame = 0; frame < frames; ++frame) {
                for (int s = 0; s < params.substeps; ++s) {
                    cpuSolver.forwardStep(mesh);
                    gpuSolver.updateInertiaFromCPU(mesh);
                    gpuSolver.solveVBD(params.numIterations, dtSqrReciprocal, params.mu, params.lambda,
                                       params.useAcceleration, params.accelerationRho);
                    gpuSolver.copyPositionsToCPU(mesh);
                    mesh.updateVelocity(subDt);
                }
                f.write((const char*)mesh.mVertPos.data(), sizeof(float) * 3 * mesh.numVerts);
            }
        } else {
            for (int frame = 0; frame < frames; ++frame) {
                cpuSolver.step(mesh);
                f.write((const char*)mesh.mVertPos.data(), sizeof(float) * 3 * mesh.numVerts);
            }
            std::cout << "energy " << cpuSolver.computeTotalEnergy(mesh) << std::endl;
        }
        if (argc > 11) mesh.exportToObj(argv[11]);
        std::cout << "done " << frames << " frames" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
