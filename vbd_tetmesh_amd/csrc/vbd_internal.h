// Internal declarations shared by the device code (vbd_kernels.hip) and the C ABI
// implementation (vbd_capi.hip).  Not installed; callers use include/vbd_capi.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "vbd_capi.h"

namespace vbd {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef int i4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

// Device image of one uploaded mesh (or batch union).  Vertex arrays are in "slot" order:
// free vertices color-major (ascending caller id inside a color), then pinned vertices.
// Per-(vertex, tet) "pair" records are laid out in the same order, so one color launch
// streams one contiguous range of every pair array.
struct DevMesh {
    int nV = 0, nT = 0, nFree = 0, nColors = 0;
    int64_t nPairs = 0;
    int maxValence = 0;
    bool perTetMat = false;
    std::vector<int> colorSlotOff;   // host copy, nColors+1: slot range of each color (free only)
    std::vector<int> colorFullOff;   // host copy, nColors+1: offsets of the full color lists
    std::vector<int> valenceHist;    // host copy: number of free vertices per valence

    f4* pos = nullptr;      // nV   current positions (w unused)
    f4* inert = nullptr;    // nV   inertia target y
    f4* vel = nullptr;      // nV   velocity
    f4* xprev = nullptr;    // nV   position at the start of the step
    f4* save[2] = {nullptr, nullptr};  // nV  x at the start of an iteration (Chebyshev ping-pong)
    float* mass = nullptr;  // nV
    int* row = nullptr;     // nFree+1 pair offsets
    int* perm = nullptr;    // nV slot -> caller vertex id
    int* inv = nullptr;     // nV caller vertex id -> slot
    i4* ptet = nullptr;     // nPairs  3 corner slots (27 bits; bit 31 of .x: the vertex is corner 0;
                            //         bits 27-30 of .x/.y/.z: owner offset from chunkBase), V0 bits in .w
    f4* pa = nullptr;       // nPairs  R0.xyz, R1.x  (three DmInv rows, see k_build_pairs; built on first use)
    f4* pb = nullptr;       // nPairs  R1.yz, R2.xy
    float* pc = nullptr;    // nPairs  R2.z
    f2* pmat = nullptr;     // nPairs  mu, lambda (per-tet materials only)
    // compact records of the tiled kernel (with ptet): F q = beta (x_s - y_a) + alpha1 (y_b - o)
    // + alpha2 (y_c - o) with o = corner 0, C q = d (y_b - y_a) x (y_c - y_a), J = C q . (x_s - y_a)
    f4* pq = nullptr;       // nPairs  beta, alpha1, alpha2, d
    // GPU-built structures kept in the caller's numbering (exposed for verification)
    int* adjOff = nullptr;    // nV+1
    int* adjEntry = nullptr;  // 4*nT, tet*4 + corner
    int* colorList = nullptr; // nV
    // per-tet arrays kept for the energy reduction
    i4* tets = nullptr;     // nT corner slots (caller tet order)
    float* dminv = nullptr; // 9*nT
    float* vol = nullptr;   // nT
    float* tmu = nullptr;   // nT or null
    float* tlam = nullptr;  // nT or null
    // tiled color kernel: owner bases of 4096-record chunks (read by the tile build) and the
    // tile cut of each color
    int* chunkBase = nullptr;          // ceil(nPairs / 4096) + 1 chunk bases
    bool tiledOk = false;              // owner offsets fit the 12 record bits
    std::vector<int> hostRow;          // host copy of row (nFree+1)
    int* tileStart = nullptr;          // per color: its tile start slots, then the color's end slot
    int* tileRec = nullptr;            // row[] at each tileStart entry (first record of the tile)
    int tileRecords = 0;               // record cap per tile of the current cut (0: none)
    int tileVerts = 0;                 // vertex cap per tile of the current cut
    int tileEntries = 0;               // entries of tileStart / tileRec (tiles + one per color)
    // the cut's records in tileRecords padded slots per tileStart entry (owner bits: the
    // tile-local vertex index); tpa/tpb/tpc only when the corotated material needs them
    i4* tpt = nullptr;
    f4* tpq = nullptr;
    f2* tpm = nullptr;
    f4* tpa = nullptr;
    f4* tpb = nullptr;
    float* tpc = nullptr;
    int64_t padBytes = 0;              // bytes of the padded records
    std::vector<int> colorTileBase;    // index of each color's first entry in tileStart
    std::vector<int> colorTileCount;   // tiles per color
    int* dColorTileBase = nullptr;     // device copies of the two (one allocation, base then count)
    int* dColorTileCount = nullptr;
    int64_t bytes = 0;
};



struct ContactArgs {
    int on;
    int axis;
    float k;
    float plane;
};

// ---- launchers (vbd_kernels.hip); all return hipError_t of the launch -----------------
bool launch_config_supported(int lpv, int unroll, int block);
hipError_t launch_color(const DevMesh& m, int color, int lpv, int unroll, int block, float sdt, float mu, float lambda,
                        const ContactArgs& ct, f4* save, hipStream_t s);
// corotated material of the reference GPU kernel (per-vertex lane groups, lpv 8/16/32)
hipError_t launch_color_corotated(const DevMesh& m, int color, int lpv, int block, float sdt, float mu, float lambda,
                                  const ContactArgs& ct, f4* save, hipStream_t s);
// tiled variant: per_thread records per thread, tiles of per_thread*block records and fewer than
// block/8 vertices (build_tiles first)
bool tiled_config_supported(int per_thread, int block);
// corotated: the reference GPU kernel's material on the full records (same tiles)
hipError_t launch_color_tiled(const DevMesh& m, int color, int per_thread, int block, float sdt, float mu,
                              float lambda, const ContactArgs& ct, f4* save, hipStream_t s, bool corotated = false);
// tile cut for the tiled kernel with (per_thread, block): records = per_thread * block,
// max_verts = block / 8 - 1
hipError_t build_tiles(DevMesh& m, int records, int max_verts, hipStream_t s, std::string& err, bool full = false);
// A tile cut (tile tables and padded records) moved out of a DevMesh and back: the autotuner
// keeps the current cut allocated while its candidates build their own, and puts it back when
// the current configuration wins (a rebuilt cut lands elsewhere in device memory, and the
// first one measured faster).
struct TileCut {
    int* tileStart = nullptr;
    int* tileRec = nullptr;
    int tileRecords = 0, tileVerts = 0, tileEntries = 0;
    i4* tpt = nullptr;
    f4* tpq = nullptr;
    f2* tpm = nullptr;
    f4* tpa = nullptr;
    f4* tpb = nullptr;
    float* tpc = nullptr;
    int64_t padBytes = 0;
    std::vector<int> colorTileBase, colorTileCount;
    int* dColorTileBase = nullptr;
    int* dColorTileCount = nullptr;
};
void detach_tiles(DevMesh& m, TileCut& out);                  // m is left without a cut
void attach_tiles(DevMesh& m, TileCut& in, hipStream_t s);     // frees m's cut, moves `in` back
void free_tile_cut(TileCut& t);
hipError_t launch_blend(const DevMesh& m, float omega, f4* pp, hipStream_t s, bool keep = false);
// One launch per sweep: every color through the tiled kernel's tile_solve with a grid barrier
// between colors, then the Chebyshev blend when omega > 1 (pp: x at the start of the previous
// sweep).  grid: a multiple of 8 not above the resident blocks (persistent_occupancy per CU x
// CUs); bar: PERSIST_BAR_WORDS zeroed words; spin: polls before a barrier gives up and raises
// bar[PERSIST_FLAG].
constexpr int PERSIST_LINE = 32;                    // words per 128-byte line
constexpr int PERSIST_BAR_WORDS = 11 * PERSIST_LINE;  // global count, generation, flag, 8 group counts
constexpr int PERSIST_FLAG = 2 * PERSIST_LINE;
int persistent_occupancy(const DevMesh& m, int per_thread, int block, const ContactArgs& ct, bool save_on);
hipError_t launch_sweep_persistent(const DevMesh& m, int per_thread, int block, int grid, float sdt, float mu,
                                   float lambda, const ContactArgs& ct, f4* save, const f4* pp, float omega,
                                   unsigned* bar, unsigned spin, hipStream_t s);
hipError_t launch_predict(const DevMesh& m, float h, const float g[3], hipStream_t s);
hipError_t launch_velocity(const DevMesh& m, float h, float damping, hipStream_t s);
hipError_t launch_copy_inertia(const DevMesh& m, hipStream_t s);
hipError_t launch_scatter_pos(int n, const float* src3, const int* inv, f4* dst, hipStream_t s);
hipError_t launch_gather_pos(int n, const f4* src, const int* inv, float* dst3, hipStream_t s);
// energies in double, fixed summation order; partial holds energy_partials(m) + SUM_STAGE
// doubles (the tail is the second stage of the final sum)
// mode 0: the reference's log-NH diagnostic, 1: the stable NH (J-1) elastic energy,
// 2: the corotated model's energy
int energy_partials(const DevMesh& m);
hipError_t launch_energy(const DevMesh& m, int mode, float mu, float lambda, double* partial, double* out,
                         hipStream_t s);
hipError_t launch_inertia_energy(const DevMesh& m, double sdt, const ContactArgs& ct, double* partial, double* out,
                                 hipStream_t s);
hipError_t launch_copy_f4(int n, const f4* src, f4* dst, hipStream_t s);
// differentiation passes (mode: 0 adjoint GS update, 1 adjoint residual, 2 A*w, 3 forward residual)
hipError_t launch_pass(int mode, const DevMesh& m, int slot0, int count, int lpv, int block, f4* w, const f4* rhs,
                       f4* out, double* r2, f4* save, float mu, float lambda, float sdt, const ContactArgs& ct,
                       hipStream_t s);
// Sum of v[0, n) into out[0] in a fixed order (bit-reproducible).  Above SUM_TWO_STAGE
// entries, with stage (SUM_STAGE doubles) given: SUM_STAGE blocks each sum one contiguous
// chunk, then one block sums the chunk sums; otherwise one block strides the whole range.
constexpr int SUM_STAGE = 256;
constexpr int SUM_TWO_STAGE = 16384;
hipError_t launch_sum_fixed(int n, const double* v, double* out, hipStream_t s, double* stage = nullptr);
hipError_t launch_norm2_f4(int n, const f4* v, double* r2, hipStream_t s);
hipError_t launch_gravity_load2(const DevMesh& m, const float g[3], double* r2, hipStream_t s);
hipError_t launch_param_grad(const DevMesh& m, const f4* w, float* dmu, float* dlam, hipStream_t s);
hipError_t launch_update_pmat(const DevMesh& m, hipStream_t s);
// the current tile cut's padded (mu, lambda) slots from pmat (after launch_update_pmat)
hipError_t launch_pad_pmat(const DevMesh& m, hipStream_t s);
hipError_t launch_blend_vec(int n, f4* v, const f4* pp, float omega, hipStream_t s);

// Device mesh build.  Host inputs are already on the device (d_* pointers, caller numbering).
// colorOf: color per vertex (host coloring); fixed: pinned flags.  Allocates every DevMesh
// array; returns a HIP error or hipSuccess, message in err.
hipError_t build_device_mesh(DevMesh& m, int nV, int nT, int nColors, const i4* d_tets, const int* d_colorOf,
                             const uint8_t* d_fixed, const float* d_dminv, const float* d_vol, const float* d_mu,
                             const float* d_lam, const float* d_x3, const float* d_y3, const float* d_v3,
                             const float* d_mass, const int* h_perm, const int* h_slot_off, hipStream_t s,
                             std::string& err);
void free_device_mesh(DevMesh& m);
// The 52-byte full pair records (pa, pb, pc) read by the per-vertex kernels, the corotated
// material and the differentiation passes: built on first use, released when no kernel of the
// handle's configuration reads them (release also frees the tile cut's padded copies).
hipError_t ensure_full_records(DevMesh& m, hipStream_t s, std::string& err);
void release_full_records(DevMesh& m, hipStream_t s);

// FP64 state of the differentiation entry points (slot order, nV double4 each): positions,
// inertia target, velocity, x at the start of the step, the Chebyshev buffers, and the adjoint
// vectors.
struct Dev64 {
    d4* pos = nullptr;
    d4* inert = nullptr;
    d4* vel = nullptr;
    d4* xprev = nullptr;
    d4* save[2] = {nullptr, nullptr};
    d4* w = nullptr;
    d4* rhs = nullptr;
    d4* out = nullptr;
    d4* wsave[2] = {nullptr, nullptr};
    // conjugate gradients (allocated on first use): residual, preconditioned residual, direction,
    // its product with A, the block-Jacobi inverse (two double4 per slot), two arrays of per-slot
    // products (2 (nV + SUM_STAGE) doubles), scalars
    d4* cg_r = nullptr;
    d4* cg_z = nullptr;
    d4* cg_p = nullptr;
    d4* cg_s = nullptr;
    d4* cg_da = nullptr;
    d4* cg_db = nullptr;
    double* cg_part = nullptr;
    double* cg_scal = nullptr;
};
// slots of Dev64::cg_scal: gamma (two, alternate iterations), delta, alpha (two), r.r, breakdown flag
constexpr int CG_G = 0, CG_D = 2, CG_A = 3, CG_RR = 5, CG_BAD = 6, CG_SCALARS = 8;
hipError_t alloc_dev64(Dev64& d, int nV);
hipError_t alloc_cg64(Dev64& d, int nV);
hipError_t launch_cg_init64(int n, const Dev64& d, double* partRz, double* partRr, hipStream_t s);
hipError_t launch_cg_update64(int n, int k, const Dev64& d, d4* x, const d4* w, double* partRz, double* partRr,
                              hipStream_t s);
hipError_t launch_sum3_fixed(int n, const double* v0, double* o0, const double* v1, double* o1, const double* v2,
                             double* o2, hipStream_t s);
hipError_t launch_scale_d4(int n, const d4* src, d4* dst, double a, hipStream_t s);
hipError_t launch_newton_move64(int n, d4* pos, const d4* base, const d4* dx, double alpha, hipStream_t s);
void free_dev64(Dev64& d);
// mode: 0 Gauss-Seidel position update (one color), 1 forward residual grad G, 2 adjoint
// block-GS update (one color), 3 adjoint residual rhs - A w, 4 A w (with r2: w . A w per slot),
// 5 the inverse diagonal blocks (out, save: two double4 per slot); on free slots
// [slot0, slot0 + count), reading the full pair records
hipError_t launch_pass64(int mode, const DevMesh& m, const Dev64& d, int slot0, int count, d4* w, const d4* rhs, d4* out,
                         double* r2, d4* save, double mu, double lambda, double sdt, const ContactArgs& ct,
                         hipStream_t s);
hipError_t launch_predict64(const DevMesh& m, const Dev64& d, double h, const float g[3], hipStream_t s);
hipError_t launch_velocity64(const DevMesh& m, const Dev64& d, double h, double damping, hipStream_t s);
hipError_t launch_blend64(int n, d4* v, const d4* pp, double omega, hipStream_t s);
hipError_t launch_copy_d4(int n, const d4* src, d4* dst, hipStream_t s);
hipError_t launch_f4_to_d4(int n, const f4* src, d4* dst, hipStream_t s);
hipError_t launch_d4_to_f4(int n, const d4* src, f4* dst, hipStream_t s);
hipError_t launch_norm2_d4(int n, const d4* v, double* r2, hipStream_t s);
hipError_t launch_scatter_d(int n, const double* src3, const int* inv, d4* dst, hipStream_t s);
hipError_t launch_scatter_fd(int n, const float* src3, const int* inv, d4* dst, hipStream_t s);
hipError_t launch_gather_d(int n, const d4* src, const int* inv, double* dst3, hipStream_t s);
hipError_t launch_gather_df(int n, const d4* src, const int* inv, float* dst3, hipStream_t s);
hipError_t launch_param_grad64(const DevMesh& m, const Dev64& d, const d4* w, float* dmu, float* dlam, hipStream_t s);
hipError_t launch_param_grad64(const DevMesh& m, const Dev64& d, const d4* w, double* dmu, double* dlam,
                               hipStream_t s);

}  // namespace vbd
