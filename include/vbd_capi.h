/*
 * vbd_capi.h -- C ABI of the MI355X (gfx950) Vertex Block Descent solver.
 *
 * Plain pointers and sizes only; no C++ or torch types cross this boundary.  Every entry
 * point returns an int status (VBD_OK = 0); vbd_last_error() returns the message of the
 * last failure on a handle.  A handle is not thread-safe; distinct handles may be used from
 * distinct host threads.  All device work of a handle is issued on that handle's own HIP
 * stream; host-buffer entry points (upload/download) synchronise that stream.
 *
 * Vertex data crossing the ABI uses the reference TetMesh layout: 3 floats per vertex, xyz
 * contiguous (TVerticesMat, Types.h:53), in the caller's vertex numbering.  Tets are 4 int32
 * per tet (TetVIdsArr, Types.h:58); DmInv is 9 floats per tet, row-major, as
 * VBD_GPU_Solver::initFromCPU flattens it (VBD_GPU_Wrapper.cu:277-282).
 *
 * Reference interfaces replaced (file:line in extinguish7/VBD_TetMesh):
 *   vbd_create / vbd_destroy      VBD_GPU_Solver ctor / dtor   VBD_GPU_Wrapper.h:9-12, .cu:373-382
 *   vbd_init                      VBD_GPU_Solver::initFromCPU  VBD_GPU_Wrapper.h:15-16, .cu:246-315
 *   vbd_solve_iteration           VBD_GPU_Solver::solveIteration  VBD_GPU_Wrapper.h:19, .cu:317-329
 *   vbd_solve                     VBD_GPU_Solver::solveVBD     VBD_GPU_Wrapper.h:28-29, .cu:332-362
 *   vbd_download_positions        VBD_GPU_Solver::copyPositionsToCPU  VBD_GPU_Wrapper.h:22, .cu:364-366
 *   vbd_upload_inertia            VBD_GPU_Solver::updateInertiaFromCPU VBD_GPU_Wrapper.h:25, .cu:369-371
 *   vbd_set_params                VBDsolver::setParams          VBD.h:41, VBD.cpp:19-21
 *   vbd_step                      VBDsolver::step               VBD.h:58, VBD.cpp:60-101
 *   vbd_forward_step              VBDsolver::forwardStep        VBD.h:99, VBD.cpp:103-112
 *   vbd_update_velocity           TetMesh::updateVelocity       TetMesh.h:305-311
 *   vbd_mesh_initialize           TetMesh::initialize (mass, adjacency)  TetMesh.h:213-275
 *   vbd_color_vertices            GraphColoring::colorVertices  GraphColoring.h:25-29, .cpp:28-66
 *   vbd_precompute_tet_data       VBDsolver::precomputeTetData  VBD.h:51, VBD.cpp:36-58
 *   vbd_accelerator_omega         VBDsolver::getAcceleratorOmega  VBD.h:139, VBD.cpp:302-311
 *   vbd_total_energy              VBDsolver::computeTotalEnergy   VBD.h:69, VBD.cpp:258-288
 * Entry points without a reference counterpart say so in their comment.
 */
#ifndef VBD_CAPI_H
#define VBD_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VBD_CAPI_VERSION 2

enum vbd_status {
    VBD_OK = 0,
    VBD_ERR_ARG = 1,     /* invalid argument (null pointer, size, index out of range) */
    VBD_ERR_HIP = 2,     /* a HIP runtime call failed; message names it */
    VBD_ERR_STATE = 3,   /* call order violated (e.g. solve before init) */
    VBD_ERR_ALLOC = 4,   /* device or host allocation failed */
    VBD_ERR_NUMERIC = 5  /* non-finite input detected */
};

typedef struct vbd_solver vbd_solver;

/* Graphs of vbd_solve calls and of the converged-step / adjoint-solve iteration blocks are cached
 * per handle, keyed by the scalars they bake in.  A key is captured on its second use (the first
 * runs the same kernels directly), and at most this many are kept (least recently used evicted),
 * so callers that vary dt or the material scalars on every call hold no growing set of graphs. */
#define VBD_MAX_CACHED_GRAPHS 32

/* Mesh as handed to vbd_init.  All arrays are borrowed for the duration of the call. */
typedef struct vbd_mesh_desc {
    int32_t num_verts;
    int32_t num_tets;
    const float* positions;     /* 3*num_verts, current positions (TetMesh::mVertPos) */
    const float* inertia;       /* 3*num_verts or NULL (TetMesh::mInertia); NULL -> positions */
    const float* velocity;      /* 3*num_verts or NULL (TetMesh::mVelocity); NULL -> zero */
    const float* mass;          /* num_verts (TetMesh::vertexMass) */
    const uint8_t* is_fixed;    /* num_verts, nonzero = pinned (TetMesh::isFixed) */
    const int32_t* tets;        /* 4*num_tets (TetMesh::tets) */
    const float* dminv;         /* 9*num_tets row-major (VBDsolver::tetDmInv) */
    const float* rest_volume;   /* num_tets (VBDsolver::tetRestVolume) */
    const float* tet_mu;        /* num_tets or NULL -> scalar mu of each solve call / params */
    const float* tet_lambda;    /* num_tets or NULL (must be NULL iff tet_mu is NULL) */
    int32_t num_colors;
    const int32_t* color_offsets; /* num_colors+1 offsets into color_verts */
    const int32_t* color_verts;   /* every vertex exactly once (VBDsolver::vertexColors) */
    int32_t slot_order;           /* device vertex order inside a color: 0 = auto (Morton order of the
                                     positions when that brings tet corners closer together than the
                                     caller's ids, else the ids), 1 = ascending vertex id, 2 = Morton.
                                     Results are identical either way; only memory locality changes. */
} vbd_mesh_desc;

/* Per-step parameters, VBDParams (VBD.h:14-25) plus build-defined extensions. */
typedef struct vbd_step_params {
    float dt;                   /* frame time step (VBDParams::dt) */
    int32_t num_iterations;     /* VBD iterations per substep */
    int32_t substeps;
    float gravity[3];
    int32_t use_acceleration;   /* Chebyshev acceleration on/off */
    float acceleration_rho;
    float mu;                   /* used when the mesh has no per-tet materials */
    float lambda;
    float velocity_damping;     /* 0.999 in the reference (TetMesh.h:308) */
    int32_t start_from_inertia; /* 0: iterate from x_t (VBD.cpp); 1: from y (VBD_GPU_Wrapper.cu:334) */
    int32_t contact_on;         /* extension: penalty plane x[axis] >= plane, E = k/2 max(0, plane-x)^2 */
    int32_t contact_axis;
    float contact_k;
    float contact_plane;
} vbd_step_params;

typedef struct vbd_info {
    int32_t num_verts;
    int32_t num_tets;
    int32_t num_free;           /* vertices that are solved (not pinned) */
    int32_t num_colors;
    int64_t num_pairs;          /* (free vertex, incident tet) records streamed per iteration */
    int32_t max_valence;
    int32_t per_tet_material;
    int32_t lanes_per_vertex;   /* current launch configuration of the color kernel (0: tiled) */
    int32_t pairs_per_lane;
    int32_t block_size;
    int64_t device_bytes;       /* bytes of device memory the handle holds */
    int32_t cached_graphs;      /* instantiated hipGraphs the handle keeps (the step's, plus at most
                                   VBD_MAX_CACHED_GRAPHS of vbd_solve / converged-step / adjoint blocks) */
    int32_t full_records;       /* 1 while the 52-byte full pair records are resident (built on first
                                   use by the corotated material, the per-vertex kernels and the
                                   differentiation entry points) */
    int32_t persistent_grid;    /* blocks of one persistent sweep launch for the step's parameters
                                   (0: one launch per color; see vbd_set_persistent) */
} vbd_info;

/* ---- lifecycle -------------------------------------------------------------------- */
int vbd_create(vbd_solver** out, int32_t device);
void vbd_destroy(vbd_solver* h);
const char* vbd_last_error(const vbd_solver* h);
/* Message of the last vbd_create failure or of a host helper (no handle). */
const char* vbd_global_last_error(void);
int32_t vbd_capi_version(void);

/* ---- upload / transfer ------------------------------------------------------------ */
/* Builds on the device: the vertex->tet CSR (entries tet*4 + corner, ascending tet id per
 * vertex), the per-color vertex lists, a color-major renumbering of the free vertices and
 * the per-(vertex, tet) records streamed by the color kernel. */
int vbd_init(vbd_solver* h, const vbd_mesh_desc* mesh);
/* Batched upload (no reference counterpart): the disjoint union of num_meshes meshes, each
 * with its own coloring, tets in its own local numbering; colors are merged by index, so one
 * color launch covers every mesh.  Vertex arrays on later transfers are the concatenation in
 * mesh order. */
int vbd_init_batch(vbd_solver* h, int32_t num_meshes, const vbd_mesh_desc* meshes);
int vbd_upload_positions(vbd_solver* h, const float* positions);
int vbd_upload_inertia(vbd_solver* h, const float* inertia);
int vbd_upload_velocities(vbd_solver* h, const float* velocity);
int vbd_download_positions(vbd_solver* h, float* positions);
int vbd_download_velocities(vbd_solver* h, float* velocity);
int vbd_download_inertia(vbd_solver* h, float* inertia);
/* Positions at the start of the last substep (TetMesh::mVertPrevPos after VBDsolver::step). */
int vbd_download_previous_positions(vbd_solver* h, float* positions);
int vbd_get_info(const vbd_solver* h, vbd_info* info);
/* GPU-built adjacency in the caller's numbering: offsets (num_verts+1), entries (4*num_tets)
 * = tet*4 + local corner. */
int vbd_get_adjacency(vbd_solver* h, int32_t* offsets, int32_t* entries);
/* GPU-built per-color vertex lists: offsets (num_colors+1), vertices (num_verts). */
int vbd_get_color_lists(vbd_solver* h, int32_t* offsets, int32_t* verts);

/* ---- solve (VBD_GPU_Solver surface) ------------------------------------------------ */
int vbd_solve_iteration(vbd_solver* h, float dt_sqr_reciprocal, float mu, float lambda);
int vbd_solve(vbd_solver* h, int32_t num_iterations, float dt_sqr_reciprocal, float mu, float lambda,
              int32_t use_acceleration, float rho, int32_t start_from_inertia);

/* ---- whole time step on the device (VBDsolver surface) ----------------------------- */
int vbd_set_params(vbd_solver* h, const vbd_step_params* params);
int vbd_forward_step(vbd_solver* h);
int vbd_update_velocity(vbd_solver* h);
/* num_frames calls of VBDsolver::step; the step is captured once in a hipGraph and
 * replayed.  Asynchronous: returns after enqueueing; transfers or vbd_synchronize wait. */
int vbd_step(vbd_solver* h, int32_t num_frames);
/* VBDsolver::step's energy record (VBD.cpp:92-94): with every > 0, each step also evaluates
 * the diagnostic energy of vbd_total_energy (materials: the step parameters, or per-tet) after
 * iterations 0, every, 2 every, ... of each substep into a device log (the reference logs at
 * every = 10); every <= 0 turns it off (the default).  Part of the captured step. */
int vbd_set_energy_log(vbd_solver* h, int32_t every);
/* The energy log of the last step: *count = substeps * ceil(iterations / every) values, of which
 * min(count, capacity) are written to out.  Waits for the handle's stream. */
int vbd_energy_log(vbd_solver* h, double* out, int32_t capacity, int32_t* count);
int vbd_synchronize(vbd_solver* h);
/* Enqueues all further work of the handle on `stream` (a hipStream_t of the handle's device,
 * e.g. a framework's current stream); NULL returns to the handle's own non-blocking stream.
 * Waits for the work already enqueued on the previous stream.  The caller keeps the stream
 * alive while the handle uses it. */
int vbd_set_stream(vbd_solver* h, void* stream);
/* Enables (1, default) or disables (0) hipGraph capture of vbd_step. */
int vbd_set_graph_enabled(vbd_solver* h, int32_t enabled);
/* Sweep launches (no reference counterpart).  0 (default): one launch per color and one per
 * Chebyshev blend, replayed from the step's hipGraph.  1: one persistent launch per sweep --
 * every color through the tiled kernel's tile code with a device-wide barrier between colors,
 * then the blend -- on 2 blocks per CU; bitwise the same positions, and slower on MI355X (a
 * cross-XCD barrier costs more than the kernel boundary it replaces: 1.7x to 2.6x the per-color
 * graph's time on the measured workloads).  8k (k >= 1): the same on at most 8k blocks (and at
 * most the occupancy API's count).  Applies to the tiled configuration of the stable
 * Neo-Hookean material (others launch per color).  A barrier whose blocks are not all resident
 * gives up after a bounded wait; the next vbd_synchronize or download then fails with
 * VBD_ERR_HIP. */
int vbd_set_persistent(vbd_solver* h, int32_t mode);

/* ---- material model ------------------------------------------------------------------ */
/* VBD_MATERIAL_STABLE_NEO_HOOKEAN (default): the CPU solver's material, VBDsolver / Energy.cpp.
 * VBD_MATERIAL_COROTATED_GPU: what the reference GPU kernel runs (VBD_GPU_Wrapper.cu:181 with
 *   VBD_DeviceMath.cuh:127-198): corotated linear elasticity with R from 5 polar iterations,
 *   H = V [2 mu |q|^2 I + lambda (R q)(R q)^T], and the wrapper's m/h^2 * 1.000001 Hessian
 *   diagonal (.cu:208).  Runs the tiled kernel on the full records when the launch
 *   configuration is tiled (lanes_per_vertex 0, the default), else the per-vertex kernel
 *   (lanes 8/16/32, else 16).  The differentiation entry points reject it. */
#define VBD_MATERIAL_STABLE_NEO_HOOKEAN 0
#define VBD_MATERIAL_COROTATED_GPU 1
int vbd_set_material_model(vbd_solver* h, int32_t model);

/* ---- launch configuration, autotuning and timing (no reference counterpart) ------- */
/* Color-kernel mapping.
 * lanes_per_vertex > 0: per-vertex kernel on the full records (F and cof F formed as in
 *   Energy.cpp): lanes_per_vertex lanes share one vertex; each lane issues the loads of
 *   pairs_per_lane incident-tet records per pass.  Supported: lanes 4/8/16 with 1-3 pairs per
 *   lane, lanes 32/64 with 1; block_size 64/128/192/256 (a multiple of lanes).
 * lanes_per_vertex == 0: tiled kernel on the compact records: one block per tile of
 *   consecutive vertices holding at most block_size * pairs_per_lane records (and fewer than
 *   block_size / 8 vertices); every thread takes pairs_per_lane records.  Supported:
 *   block_size 64-512 (a multiple of 64), pairs_per_lane 1 or 2, block_size * pairs_per_lane
 *   >= max valence.  Results are bitwise identical across tiled configurations.
 * The default after vbd_init is the tiled kernel with block 256 and 1 pair per lane (block 512
 * when the mean valence is 20 or more) when the maximum valence allows it. */
int vbd_set_launch_config(vbd_solver* h, int32_t lanes_per_vertex, int32_t block_size, int32_t pairs_per_lane);
/* Autotuner keyed on the valence histogram: on the first mesh with a given key it times each
 * candidate mapping over `iterations` sweeps launched back to back, restores the positions and
 * keeps the fastest, or the configuration set before the call when no candidate beats it by
 * 1% or more; later meshes with the same key reuse the choice without timing (*best_ms = -1). */
int vbd_autotune(vbd_solver* h, int32_t iterations, int32_t* best_lanes, int32_t* best_unroll, int32_t* best_block,
                 float* best_ms);
/* Runs `iterations` plain sweeps (+ blends if use_acceleration in params) with HIP events
 * around every launch on the handle's stream.  Outputs: total color-kernel milliseconds,
 * number of color launches, total blend milliseconds, number of blend launches.  Moves the
 * positions. */
int vbd_time_iterations(vbd_solver* h, int32_t iterations, double* color_ms, int32_t* color_launches,
                        double* blend_ms, int32_t* blend_launches);
/* The color launches of `sweeps` sweeps (the step's own kernels: with Chebyshev in the
 * parameters each launch also saves the start-of-sweep positions, as in vbd_step), captured in
 * one hipGraph and replayed once untimed and once between two HIP events on the handle's stream
 * (no reference counterpart).  Outputs: milliseconds per color launch (the replay's time over
 * its launches, the gaps between graph nodes included) and the number of launches timed.  Moves
 * the positions. */
int vbd_time_color_launches(vbd_solver* h, int32_t sweeps, double* ms_per_launch, int32_t* launches);
/* Bytes one sweep's color launches must move in the current configuration's layout, at least
 * (no reference counterpart): per (free vertex, incident tet) record the record as stored (32 B
 * compact for the tiled stable Neo-Hookean kernel, 52 B full for the per-vertex kernels and the
 * corotated material; +8 B with per-tet mu, lambda), per free vertex 56 B (row offset 4, x read
 * 16, inertia 16, mass 4, x write 16); the blend launches 64 B per free vertex (x and x_pp read,
 * x written, and written again as the next sweep's start).  That is a sweep of a Chebyshev step
 * after its first two, whose color launches also save the sweep's start (16 B per free vertex).
 * Tile padding and neighbour-position gathers (cache-resident) are not counted. */
int vbd_layout_bytes(const vbd_solver* h, double* color_bytes_per_sweep, double* blend_bytes_per_sweep);
/* Algorithmic bytes of one sweep, a fixed yardstick independent of the device layout:
 * color launches 45 B per free vertex (x read 12, x write 12, inertia 12, mass 4, CSR row 4,
 * fixed flag 1) + 60 B per (vertex, incident tet) (CSR entry 4, tet ids 16, DmInv 36, V0 4;
 * +8 B with per-tet mu, lambda); blend launches 36 B per free vertex.  Neighbour-position
 * gathers are not counted (cache-resident). */
int vbd_algorithmic_bytes(const vbd_solver* h, double* color_bytes_per_sweep, double* blend_bytes_per_sweep);

/* ---- host helpers (CPU; no device work) ------------------------------------------- */
/* TetMesh::initialize: per-tet rest volume, lumped mass rho*V/4 per corner and vertex->tet
 * adjacency in ascending tet order (adj_offsets num_verts+1, adj_tets 4*num_tets). */
int vbd_mesh_initialize(int32_t num_verts, const float* rest, int32_t num_tets, const int32_t* tets, float density,
                        float* mass, float* tet_rest_volume, int32_t* adj_offsets, int32_t* adj_tets);
/* Greedy first-fit coloring in ascending vertex order.  color_offsets must hold
 * num_verts+1 entries; returns the color count in *num_colors. */
int vbd_color_vertices(int32_t num_verts, const int32_t* tets, const int32_t* adj_offsets, const int32_t* adj_tets,
                       int32_t* color_of, int32_t* color_offsets, int32_t* color_verts, int32_t* num_colors);
int vbd_precompute_tet_data(const float* rest, int32_t num_tets, const int32_t* tets, float* dminv,
                            float* rest_volume);
float vbd_accelerator_omega(int32_t order, float rho, float prev_omega);
/* VBDsolver::computeTotalEnergy (VBD.cpp:258-288): device reduction of the diagnostic
 * log-Neo-Hookean energy (|det F| for the singular value product) over the current positions,
 * in double with a fixed summation order (bit-reproducible). */
int vbd_total_energy(vbd_solver* h, float mu, float lambda, double* energy);
/* No reference counterpart: for callers that cannot allocate (code under hipGraph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The energy of vbd_total_energy, same kernels
 * and summation order (bitwise the same value), computed without any device or host allocation
 * and without waiting: the per-block partial sums live in one library-wide device array of
 * VBD_NOALLOC_SCRATCH_DOUBLES doubles per GPU (1 MiB, held until the process ends), and the
 * energy is written on the handle's stream to energy_out, a device-accessible double (device
 * memory or pinned host memory) that the caller reads after vbd_synchronize or an event.  Every
 * partial sum is written before it is read, so nothing is cleared between calls.  Meshes whose
 * partial sums do not fit (more than 33,488,896 tets or free vertices) are rejected with
 * VBD_ERR_ARG.  All calls on one GPU share the array: calls that overlap on one GPU (other
 * handles, other streams) race. */
#define VBD_NOALLOC_SCRATCH_DOUBLES 131072
int vbd_total_energy_noalloc(vbd_solver* h, float mu, float lambda, double* energy_out);
/* The incremental potential the step minimises (no reference counterpart; the consistent
 * objective for convergence monitoring): G = sum_free m/(2h^2) |x - y|^2 + sum_tets V0 Psi(F)
 * (+ the penalty contact energy), Psi the stable Neo-Hookean (J-1) energy of Energy.cpp:93-97
 * (or, with VBD_MATERIAL_COROTATED_GPU, mu |F - R|^2 + lambda/2 tr(R^T F - I)^2), h = dt /
 * substeps and materials from the parameters (or per-tet).  elastic = the Psi sum. */
int vbd_incremental_potential(vbd_solver* h, double* potential, double* elastic);
/* Times `reps` back-to-back evaluations of an energy reduction with HIP events on the
 * handle's stream (no host synchronisation in between; no reference counterpart).
 * mode 0: the log-Neo-Hookean diagnostic of vbd_total_energy (mu, lambda from the step
 * parameters, 1e5 / 4e5 before vbd_set_params); mode 1: the incremental potential of
 * vbd_incremental_potential (needs vbd_set_params).  Outputs: milliseconds per evaluation and
 * its algorithmic bytes -- per tet 56 (ids 16, DmInv 36, V0 4; +8 with per-tet mu, lambda),
 * 12 per vertex (positions, read once; the per-tet corner gathers are cache-resident and not
 * counted, as for the color launches), plus 28 per free vertex in mode 1 (x, inertia target,
 * mass). */
int vbd_time_energy(vbd_solver* h, int32_t mode, int32_t reps, double* ms_per_eval, double* bytes_per_eval);

/* ---- differentiation (no reference counterpart: the reference is not differentiable) ----
 * The adjoint of one implicit step x_{n+1} = argmin G needs A = M/h^2 + K(x_{n+1}) (K = Hessian
 * of the elastic energy plus the contact penalty term).  Vectors are 3 floats per vertex in the
 * caller's numbering; entries of pinned vertices are ignored and returned as zero. */
/* Replace the per-tet materials of a mesh uploaded with tet_mu/tet_lambda. */
int vbd_set_materials(vbd_solver* h, const float* tet_mu, const float* tet_lambda);
/* |grad G| at the current positions and inertia target, and |M g| (the gravity load). */
int vbd_residual_norm(vbd_solver* h, double* norm, double* gravity_load);
/* One time step (substeps == 1) iterated until |grad G| <= max(abs_tol, rel_tol |M g|), tested
 * every check_every iterations, or until the best |grad G| has not improved by 1% for
 * stall_checks consecutive tests (0 disables; the float32 floor of |grad G|), or max_iterations;
 * then the velocity update.  Deterministic: the same inputs give bit-identical results (no
 * atomics in any reduction). */
int vbd_step_converged(vbd_solver* h, int32_t max_iterations, int32_t check_every, float abs_tol, float rel_tol,
                       int32_t stall_checks, int32_t* iterations, double* final_norm);
/* out = A w at the current positions (rows of free vertices). */
int vbd_hessian_vector(vbd_solver* h, const float* w, float* out);
/* Solves A w = rhs by colored block Gauss-Seidel (the forward colors, exact 3x3 diagonal blocks)
 * with Chebyshev acceleration when 0 < rho < 1.  w is the initial guess on entry and the
 * solution on exit; stops when |rhs - A w| <= rel_tol |rhs| (tested every check_every sweeps)
 * or on stagnation as in vbd_step_converged.  Returns VBD_ERR_NUMERIC, w unchanged, when the
 * residual grows past 1e3 x max(|rhs|, first residual) or is non-finite (rho too large; plain
 * Gauss-Seidel, rho = 0, converges for the SPD A). */
int vbd_adjoint_solve(vbd_solver* h, const float* rhs, float* w, int32_t max_iterations, int32_t check_every,
                      float rel_tol, float rho, int32_t stall_checks, int32_t* iterations, double* rel_residual);
/* Per-tet -(d^2 E / dx dmu)^T w and -(d^2 E / dx dlambda)^T w at the current positions:
 * dmu_t = -V (F - C) : dF(w), dlambda_t = -V (J - 1) C : dF(w), in the caller's tet order. */
int vbd_parameter_gradients(vbd_solver* h, const float* w, float* dmu, float* dlambda);
/* Precision of the differentiation entry points above (no reference counterpart).  32, the
 * default: FP32 kernels on the FP32 state; a converged step's |grad G| stalls near H ulp(x) per
 * vertex in absolute coordinates (about 5e-3 |M g| on a 16^3 unit cube).  64: the handle keeps an
 * FP64 copy of the state (positions, velocities, inertia target, step-start positions), and
 * vbd_step_converged, vbd_residual_norm, vbd_hessian_vector, vbd_adjoint_solve and
 * vbd_parameter_gradients run FP64 per-vertex kernels on it (the pair records' FP32 DmInv rows
 * promoted exactly), so a converged step reaches |grad G| below 1e-6; after each converged step
 * the FP32 arrays hold its rounding.  Any entry point that writes the FP32 state (uploads, the
 * FP32 steps and solves) makes the next FP64 call start from the FP32 state. */
int vbd_set_differentiation_precision(vbd_solver* h, int32_t bits);
/* The FP64 state's positions and velocities, 3 doubles per vertex in the caller's numbering
 * (precision 64 only; for checkpoints that must not round). */
int vbd_download_state_f64(vbd_solver* h, double* positions, double* velocities);
/* Sets the FP64 state's positions and/or velocities (either may be NULL: kept) and their FP32
 * rounding (precision 64 only). */
int vbd_upload_state_f64(vbd_solver* h, const double* positions, const double* velocities);
/* vbd_hessian_vector, vbd_adjoint_solve and vbd_parameter_gradients with double vectors and
 * outputs (precision 64 only): a rollout's backward recursion (a <- M/h^2 w - (c/h) b) cancels
 * most of its terms, so FP32 vectors there would cost the gradients about 1e-4 of their size. */
int vbd_hessian_vector_f64(vbd_solver* h, const double* w, double* out);
int vbd_adjoint_solve_f64(vbd_solver* h, const double* rhs, double* w, int32_t max_iterations, int32_t check_every,
                          float rel_tol, float rho, int32_t stall_checks, int32_t* iterations, double* rel_residual);
int vbd_parameter_gradients_f64(vbd_solver* h, const double* w, double* dmu, double* dlambda);
/* The adjoint system (M/h^2 + K(x)) w = rhs of vbd_adjoint_solve_f64, solved by conjugate
 * gradients with the block-Jacobi preconditioner (the inverse 3x3 diagonal blocks) instead of
 * colored block Gauss-Seidel: no counterpart in the reference (it has no adjoint).  w holds the
 * starting guess on entry; the solve stops when |rhs - A w| <= rel_tol |rhs| (tested every
 * check_every iterations) or after max_iterations.  VBD_ERR_NUMERIC, with w unchanged, when an
 * iteration finds p.Ap or r.z not positive (A not positive definite at this state) or the
 * residual grows 1e3-fold.  Precision 64 only. */
int vbd_adjoint_solve_cg_f64(vbd_solver* h, const double* rhs, double* w, int32_t max_iterations,
                             int32_t check_every, float rel_tol, int32_t* iterations, double* rel_residual);

/* Method of vbd_step_converged (no counterpart in the reference, whose step runs a fixed number
 * of VBD iterations): VBD_CONVERGED_VBD (default) iterates VBD sweeps with Chebyshev;
 * VBD_CONVERGED_NEWTON (precision 64 only) takes Newton steps on the incremental potential G,
 * each solving (M/h^2 + K) dx = -grad G by block-Jacobi preconditioned conjugate gradients to
 * 1e-2 |grad G| and halving the step (at most 4 times) until |grad G| falls; a breakdown, a
 * failed line search or 20 steps hand the state to the VBD sweeps.  The converged state is the
 * same minimiser of G to the step's tolerance.  With Newton, vbd_step_converged's iterations
 * output counts Newton steps plus any VBD sweeps. */
#define VBD_CONVERGED_VBD 0
#define VBD_CONVERGED_NEWTON 1
int vbd_set_converged_method(vbd_solver* h, int32_t method);
/* Work of the last vbd_step_converged: Newton steps, conjugate-gradient iterations inside them,
 * VBD sweeps (any of the pointers may be NULL). */
int vbd_last_step_stats(const vbd_solver* h, int32_t* newton_steps, int32_t* inner_iterations, int32_t* sweeps);

#ifdef __cplusplus
}
#endif

#endif /* VBD_CAPI_H */
