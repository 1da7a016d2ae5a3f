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
// The hybrid mode also prints the set-up times (load, coloring, tet data, initFromCPU) and,
// over the frames after the first two, the host part (forwardStep + updateVelocity) and the device
// part (updateInertiaFromCPU + solveVBD + copyPositionsToCPU) per frame, as one "timing" line.
//        vbd_dropin_demo main <rest.node> <current.node> <model.ele> <frames> <material 0|1> <out.bin>
// The second form is reference main.cpp:171-232 as written: steel in mm (E = 200e3, nu = 0.3,
// density 7800e-12), a warm start from the current-shape node file, the y <= min + 5% pin on the
// current positions, dt = 1e-3 with 2 substeps, 50 iterations, Chebyshev rho 0.75, no gravity,
// and the hybrid frame loop around solveVBD; material 1 is the reference GPU kernel's corotated
// model, 0 the CPU solver's stable Neo-Hookean.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "vbd/TetGenIO.h"
#include "vbd/VBD.h"
#include "vbd/VBD_GPU_Wrapper.h"

using namespace VBD;

// main.cpp:171-232: returns the process status
static int reference_main(int argc, char** argv) {
    if (argc < 8) {
        std::cerr << "usage: vbd_dropin_demo main <rest.node> <current.node> <model.ele> <frames> <material> <out.bin>\n";
        return 2;
    }
    const int frames = std::atoi(argv[5]), material = std::atoi(argv[6]);
    try {
        Real youngs_modulus = 200.0e3f;
        Real poisson_ratio = 0.3f;
        Real mu = youngs_modulus / (2.0f * (1.0f + poisson_ratio));
        Real lambda = youngs_modulus * poisson_ratio / ((1.0f + poisson_ratio) * (1.0f - 2.0f * poisson_ratio));
        Real density = 7800.0e-12f;
        TetMesh mesh;
        if (!loadTetgenModel(argv[2], argv[3], argv[4], mesh, density, mu, lambda, 1, 0.05f)) {
            std::cerr << "failed to load " << argv[2] << "\n";
            return 1;
        }
        VBDParams params;
        params.dt = 0.001f;
        params.numIterations = 50;
        params.substeps = 2;
        params.gravity = Vec3(0.0f, 0.0f, 0.0f);
        params.useAcceleration = true;
        params.accelerationRho = 0.75f;
        params.numThreads = 4;
        params.mu = mu;
        params.lambda = lambda;
        VBDsolver cpuSolver;
        cpuSolver.setParams(params);
        cpuSolver.initializeColoring(mesh);
        cpuSolver.precomputeTetData(mesh);
        VBD_GPU_Solver gpuSolver;
        gpuSolver.setMaterialModel(material);
        gpuSolver.initFromCPU(mesh, cpuSolver.getVertexColors(), cpuSolver.getTetDmInv(), cpuSolver.getTetRestVol());
        Real subDt = params.dt / params.substeps;
        Real dtSqrReciprocal = 1.0f / (subDt * subDt);
        std::cout << "mesh " << mesh.numVerts << " verts, " << mesh.numTets << " tets, "
                  << cpuSolver.getVertexColors().size() << " colors, mu " << mu << " lambda " << lambda << std::endl;
        std::ofstream f(argv[7], std::ios::binary);
        for (int frame = 0; frame < frames; ++frame) {
            for (int step = 0; step < params.substeps; ++step) {
                cpuSolver.forwardStep(mesh);
                gpuSolver.updateInertiaFromCPU(mesh);
                gpuSolver.solveVBD(params.numIterations, dtSqrReciprocal, params.mu, params.lambda,
                                   params.useAcceleration, params.accelerationRho);
                gpuSolver.copyPositionsToCPU(mesh);
                mesh.updateVelocity(subDt);
            }
            f.write((const char*)mesh.mVertPos.data(), sizeof(float) * 3 * mesh.numVerts);
        }
        std::cout << "done " << frames << " frames" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc > 1 && std::string(argv[1]) == "main") return reference_main(argc, argv);
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
    using clk = std::chrono::steady_clock;
    auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
    try {
        const auto tl0 = clk::now();
        TetMesh mesh;
        if (!loadTetgenModel(prefix + ".node", "", prefix + ".ele", mesh, 1000.0f, mu, lambda, pinAxis, pinFrac)) {
            std::cerr << "failed to load " << prefix << "\n";
            return 1;
        }
        const auto tl1 = clk::now();
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
        const auto tc0 = clk::now();
        cpuSolver.initializeColoring(mesh);
        const auto tc1 = clk::now();
        cpuSolver.precomputeTetData(mesh);
        const auto tc2 = clk::now();
        std::cout << "mesh " << mesh.numVerts << " verts, " << mesh.numTets << " tets, "
                  << cpuSolver.getVertexColors().size() << " colors" << std::endl;
        if (mode == "hybrid") {
            VBD_GPU_Solver gpuSolver;
            gpuSolver.initFromCPU(mesh, cpuSolver.getVertexColors(), cpuSolver.getTetDmInv(), cpuSolver.getTetRestVol());
            const auto tc3 = clk::now();
            const Real subDt = params.dt / params.substeps;
            const Real dtSqrReciprocal = 1.0f / (subDt * subDt);
            double host = 0.0, device = 0.0;
            for (int frame = 0; frame < frames; ++frame) {
                for (int s = 0; s < params.substeps; ++s) {
                    const auto t0 = clk::now();
                    cpuSolver.forwardStep(mesh);
                    const auto t1 = clk::now();
                    gpuSolver.updateInertiaFromCPU(mesh);
                    gpuSolver.solveVBD(params.numIterations, dtSqrReciprocal, params.mu, params.lambda,
                                       params.useAcceleration, params.accelerationRho);
                    gpuSolver.copyPositionsToCPU(mesh);
                    const auto t2 = clk::now();
                    mesh.updateVelocity(subDt);
                    const auto t3 = clk::now();
                    if (frame > 1) {  // the first call runs eagerly, the second captures the solve's graph
                        host += secs(t0, t1) + secs(t2, t3);
                        device += secs(t1, t2);
                    }
                }
                f.write((const char*)mesh.mVertPos.data(), sizeof(float) * 3 * mesh.numVerts);
            }
            const int timed = frames > 2 ? frames - 2 : 1;
            std::printf("timing frames %d host_ms_per_frame %.4f device_ms_per_frame %.4f load_s %.3f "
                        "coloring_s %.3f tet_data_s %.3f init_from_cpu_s %.3f\n",
                        frames > 2 ? frames - 2 : 0, 1e3 * host / timed, 1e3 * device / timed, secs(tl0, tl1), secs(tc0, tc1),
                        secs(tc1, tc2), secs(tc2, tc3));
            std::fflush(stdout);
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
