// VBD CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++ restatement of the reference CPU Vertex Block Descent path, used by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
// The product library never links, loads or calls anything in this file.
//
// What it restates (file:line into the reference, extinguish7/VBD_TetMesh):
//   * TetMesh::initialize mass + adjacency ............ TetMesh.h:213-275, :291-298
//   * GraphColoring::colorVertices (greedy first-fit) .. GraphColoring.cpp:7-66
//   * VBDsolver::precomputeTetData (DmInv, V0) ........ VBD.cpp:36-58, Energy.cpp:6-14, :24-32
//   * VBDsolver::step / forwardStep / solve ........... VBD.cpp:60-158
//   * VBDsolver::solveVertex / computeTetContribution . VBD.cpp:160-256
//   * ElasticEnergy::compute_Hessian_and_Gradient ..... Energy.cpp:61-158
//   * AnalyticSolver3x3::solve ........................ Types.h:85-112
//   * applyAcceleration / getAcceleratorOmega ......... VBD.cpp:290-311
//   * TetMesh::updateVelocity/applyGravity/computeInertia/savePreviousPosition
//                                                       TetMesh.h:305-348
//   * computeTotalEnergy (diagnostic) ................. VBD.cpp:258-288, NeoHookean.cpp:6-20
//
// Arithmetic follows the reference's operation order where the reference source fixes it
// (explicit formulas, double-literal promotions such as `lambda * (J - 1.0)` evaluated in
// double then rounded to float).  Where the order is decided inside Eigen (3-term dot and
// 3x3 product sums, 3x3 inverse) this file uses a0 + (a1 + a2), Eigen's unrolled reduction
// order for length-3 reductions; Eigen is not present in this image, so that choice is
// parity-unpinned at the last-ulp level.  Compile with -ffp-contract=off.
//
// Force backends (all give the same per-vertex 3x3 block in exact arithmetic):
//   mode 0 "closed":   H_kk = V [mu |q|^2 I + lambda (Cq)(Cq)^T], g_k = V P q
//   mode 1 "faithful": 9x9 dP/dF by nine directional derivatives (Energy.cpp:115-136),
//                      only the (k,k) 3x3 block of the 12x12 chain rule is evaluated.
//   mode 2 "faithful, reference cost": as mode 1 but the full 12x12 is assembled,
//                      a std::vector is allocated per (vertex, tet) call (VBD.cpp:220) and the
//                      log-Neo-Hookean energy is evaluated every 10 iterations (VBD.cpp:93-95).
//                      Positions are bit-identical to mode 1; this mode exists for timing.
//
// Build-defined extensions (no counterpart in the reference solver):
//   * per-tet mu/lambda (TetMesh.h:31-32 stores them; the reference solver ignores them)
//   * penalty ground contact: E_c = k/2 max(0, z_p - x_a)^2 on axis a
//   * start-from-inertia flag: VBD_GPU_Solver::solveVBD copies y into x before iterating
//     (VBD_GPU_Wrapper.cu:334), VBDsolver::step does not (VBD.cpp:72-81).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

template <class R> struct M3 {
    R m[3][3];
    R& operator()(int r, int c) { return m[r][c]; }
    const R& operator()(int r, int c) const { return m[r][c]; }
};

template <class R> static inline R sum3(R a, R b, R c) { return a + (b + c); }

// Energy.cpp:6-14 (Dm / Ds as columns x_j - x_0)
template <class R>
static inline void compute_Dm(M3<R>& D, const R* x, const int32_t* tet) {
    for (int j = 0; j < 3; ++j)
        for (int c = 0; c < 3; ++c) D(c, j) = x[3 * tet[j + 1] + c] - x[3 * tet[0] + c];
}

// Energy.cpp:16-22 (F = Ds * DmInv)
template <class R>
static inline void matmul(M3<R>& out, const M3<R>& a, const M3<R>& b) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out(r, c) = sum3(a(r, 0) * b(0, c), a(r, 1) * b(1, c), a(r, 2) * b(2, c));
}

// Eigen's 3x3 determinant (row-0 expansion); same formula as Types.h:87-89
template <class R>
static inline R det3(const M3<R>& A) {
    return A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) - A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) +
           A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
}

// Energy.cpp:61-73
template <class R>
static inline M3<R> cofactor(const M3<R>& F) {
    M3<R> C;
    C(0, 0) = F(1, 1) * F(2, 2) - F(1, 2) * F(2, 1);
    C(0, 1) = F(1, 2) * F(2, 0) - F(1, 0) * F(2, 2);
    C(0, 2) = F(1, 0) * F(2, 1) - F(1, 1) * F(2, 0);
    C(1, 0) = F(2, 1) * F(0, 2) - F(2, 2) * F(0, 1);
    C(1, 1) = F(2, 2) * F(0, 0) - F(2, 0) * F(0, 2);
    C(1, 2) = F(2, 0) * F(0, 1) - F(2, 1) * F(0, 0);
    C(2, 0) = F(0, 1) * F(1, 2) - F(0, 2) * F(1, 1);
    C(2, 1) = F(0, 2) * F(1, 0) - F(0, 0) * F(1, 2);
    C(2, 2) = F(0, 0) * F(1, 1) - F(0, 1) * F(1, 0);
    return C;
}

// Energy.cpp:76-88
template <class R>
static inline M3<R> cofactor_dir(const M3<R>& F, const M3<R>& dF) {
    M3<R> dC;
    dC(0, 0) = dF(1, 1) * F(2, 2) + F(1, 1) * dF(2, 2) - dF(1, 2) * F(2, 1) - F(1, 2) * dF(2, 1);
    dC(0, 1) = dF(1, 2) * F(2, 0) + F(1, 2) * dF(2, 0) - dF(1, 0) * F(2, 2) - F(1, 0) * dF(2, 2);
    dC(0, 2) = dF(1, 0) * F(2, 1) + F(1, 0) * dF(2, 1) - dF(1, 1) * F(2, 0) - F(1, 1) * dF(2, 0);
    dC(1, 0) = dF(2, 1) * F(0, 2) + F(2, 1) * dF(0, 2) - dF(2, 2) * F(0, 1) - F(2, 2) * dF(0, 1);
    dC(1, 1) = dF(2, 2) * F(0, 0) + F(2, 2) * dF(0, 0) - dF(2, 0) * F(0, 2) - F(2, 0) * dF(0, 2);
    dC(1, 2) = dF(2, 0) * F(0, 1) + F(2, 0) * dF(0, 1) - dF(2, 1) * F(0, 0) - F(2, 1) * dF(0, 0);
    dC(2, 0) = dF(0, 1) * F(1, 2) + F(0, 1) * dF(1, 2) - dF(0, 2) * F(1, 1) - F(0, 2) * dF(1, 1);
    dC(2, 1) = dF(0, 2) * F(1, 0) + F(0, 2) * dF(1, 0) - dF(0, 0) * F(1, 2) - F(0, 0) * dF(1, 2);
    dC(2, 2) = dF(0, 0) * F(1, 1) + F(0, 0) * dF(1, 1) - dF(0, 1) * F(1, 0) - F(0, 1) * dF(1, 0);
    return dC;
}

// lambda * (J - 1.0): float * double in the reference source, so the scalar is formed in
// double and rounded once to the working type (Energy.cpp:97, :127).
template <class R>
static inline R lam_jm1(R lambda, R J) {
    return (R)((double)lambda * ((double)J - 1.0));
}

// 3x3 inverse by adjugate / determinant (Eigen's compute_inverse_size3 structure:
// cofactor column 0, det = cofactors . col0, then scaled adjugate).
template <class R>
static inline M3<R> inverse3(const M3<R>& A) {
    auto cof = [&](int i, int j) {
        int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return A(i1, j1) * A(i2, j2) - A(i1, j2) * A(i2, j1);
    };
    R c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
    R det = sum3(c00 * A(0, 0), c10 * A(1, 0), c20 * A(2, 0));
    R inv = R(1) / det;
    M3<R> out;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out(j, i) = cof(i, j) * inv;  // out = adj / det, adj = cof^T
    return out;
}

// ---------------------------------------------------------------------------------------
// Per-(tet, corner) block: gradient g_k (3) and diagonal Hessian block H_kk (3x3)
// ---------------------------------------------------------------------------------------

// Q = Ps * DmInv (Energy.cpp:101-106): row 0 = -(row0+row1+row2) of DmInv, row n = row n-1.
template <class R>
static inline void shape_row(R q[3], const M3<R>& DmInv, int k) {
    if (k == 0) {
        for (int b = 0; b < 3; ++b) q[b] = sum3(-DmInv(0, b), -DmInv(1, b), -DmInv(2, b));
    } else {
        for (int b = 0; b < 3; ++b) q[b] = DmInv(k - 1, b);
    }
}

// Closed form of the (k,k) block (exact in real arithmetic, derived from Energy.cpp:90-158):
//   H_kk = V [mu |q|^2 I + lambda (C q)(C q)^T],  g_k = V P q.
template <class R>
static inline void block_closed(const M3<R>& F, const M3<R>& DmInv, R mu, R lambda, R V, int k, R H[3][3], R g[3]) {
    R J = det3(F);
    M3<R> C = cofactor(F);
    R a = lam_jm1(lambda, J);
    M3<R> P;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) P(r, c) = mu * (F(r, c) - C(r, c)) + a * C(r, c);
    R q[3];
    shape_row(q, DmInv, k);
    R qq = sum3(q[0] * q[0], q[1] * q[1], q[2] * q[2]);
    R Cq[3];
    for (int r = 0; r < 3; ++r) {
        g[r] = V * sum3(P(r, 0) * q[0], P(r, 1) * q[1], P(r, 2) * q[2]);
        Cq[r] = sum3(C(r, 0) * q[0], C(r, 1) * q[1], C(r, 2) * q[2]);
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) H[r][c] = V * ((r == c ? mu * qq : R(0)) + lambda * Cq[r] * Cq[c]);
}

// Corotated linear elasticity of the reference GPU kernel (VBD_DeviceMath.cuh:127-198):
// R from 5 steps of R <- (R + cof(R)/det R) / 2 (stopping if |det R| < 1e-8),
// P = 2 mu (F - R) + lambda tr(R^T F - I) R,  g_k = V P q_k,
// H_kk = V [2 mu |q_k|^2 I + lambda (R q_k)(R q_k)^T] (dR/dF dropped, always SPD).
template <class R>
static inline M3<R> polar_rotation(const M3<R>& F) {
    M3<R> Rm = F;
    for (int it = 0; it < 5; ++it) {
        R d = det3(Rm);
        if (std::abs((double)d) < 1e-8) break;
        M3<R> C = cofactor(Rm);
        R inv = R(1) / d;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) Rm(r, c) = R(0.5) * (Rm(r, c) + C(r, c) * inv);
    }
    return Rm;
}

template <class R>
static inline void block_corotated(const M3<R>& F, const M3<R>& DmInv, R mu, R lambda, R V, int k, R H[3][3],
                                   R g[3]) {
    M3<R> Rm = polar_rotation(F);
    R tr = 0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) tr += Rm(r, c) * F(r, c);
    tr -= R(3);
    M3<R> P;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) P(r, c) = R(2) * mu * (F(r, c) - Rm(r, c)) + lambda * tr * Rm(r, c);
    R q[3];  // Qn as the kernel forms it: (-D0 - D1) - D2 for corner 0
    for (int b = 0; b < 3; ++b) q[b] = k == 0 ? -DmInv(0, b) - DmInv(1, b) - DmInv(2, b) : DmInv(k - 1, b);
    for (int r = 0; r < 3; ++r) g[r] = V * (P(r, 0) * q[0] + P(r, 1) * q[1] + P(r, 2) * q[2]);
    R qq = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    R rq[3];
    for (int r = 0; r < 3; ++r) rq[r] = Rm(r, 0) * q[0] + Rm(r, 1) * q[1] + Rm(r, 2) * q[2];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) H[r][c] = V * ((r == c ? R(2) * mu * qq : R(0)) + lambda * rq[r] * rq[c]);
}

// Energy.cpp:114-136: 9x9 dP/dF, column i + 3j is dP for dF = e_i e_j^T, flattened as vi + 3 vj.
template <class R>
static inline void build_HF(R HF[9][9], const M3<R>& F, const M3<R>& C, R J, R mu, R lambda) {
    R a = lam_jm1(lambda, J);
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            int col = i + 3 * j;
            M3<R> dF;
            std::memset(&dF, 0, sizeof(dF));
            dF(i, j) = R(1);
            R dJ = C(i, j);  // (C.array() * dF.array()).sum() is exactly C(i,j)
            M3<R> dC = cofactor_dir(F, dF);
            R ld = lambda * dJ;
            for (int vj = 0; vj < 3; ++vj)
                for (int vi = 0; vi < 3; ++vi)
                    HF[vi + 3 * vj][col] = (mu * (dF(vi, vj) - dC(vi, vj)) + ld * C(vi, vj)) + a * dC(vi, vj);
        }
}

// Faithful evaluation of gradient + (k,k) block (Energy.cpp:90-158 restricted to the block
// VBD.cpp:251-255 extracts; each 12x12 entry is computed independently, so the block
// entries are bit-identical to the full assembly).
template <class R>
static inline void block_faithful(const M3<R>& F, const M3<R>& DmInv, R mu, R lambda, R V, int k, R H[3][3], R g[3]) {
    R J = det3(F);
    M3<R> C = cofactor(F);
    R a = lam_jm1(lambda, J);
    M3<R> P;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) P(r, c) = mu * (F(r, c) - C(r, c)) + a * C(r, c);
    R Q[4][3];
    for (int n = 0; n < 4; ++n) shape_row(Q[n], DmInv, n);
    // g_nodes = (volume * P) * Q^T
    for (int c = 0; c < 3; ++c) g[c] = sum3((V * P(c, 0)) * Q[k][0], (V * P(c, 1)) * Q[k][1], (V * P(c, 2)) * Q[k][2]);
    R HF[9][9];
    build_HF(HF, F, C, J, mu, lambda);
    for (int c = 0; c < 3; ++c)
        for (int cp = 0; cp < 3; ++cp) {
            R val = R(0);
            for (int b = 0; b < 3; ++b)
                for (int d = 0; d < 3; ++d) val += HF[c + 3 * b][cp + 3 * d] * Q[k][b] * Q[k][d];
            H[c][cp] = V * val;
        }
}

// Full 12-gradient and 12x12 Hessian, Energy.cpp:90-158.
template <class R>
static inline void full_faithful(const M3<R>& F, const M3<R>& DmInv, R mu, R lambda, R V, R Hout[12][12], R gout[12]) {
    R J = det3(F);
    M3<R> C = cofactor(F);
    R a = lam_jm1(lambda, J);
    M3<R> P;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) P(r, c) = mu * (F(r, c) - C(r, c)) + a * C(r, c);
    R Q[4][3];
    for (int n = 0; n < 4; ++n) shape_row(Q[n], DmInv, n);
    for (int n = 0; n < 4; ++n)
        for (int c = 0; c < 3; ++c)
            gout[3 * n + c] = sum3((V * P(c, 0)) * Q[n][0], (V * P(c, 1)) * Q[n][1], (V * P(c, 2)) * Q[n][2]);
    R HF[9][9];
    build_HF(HF, F, C, J, mu, lambda);
    for (int n = 0; n < 4; ++n)
        for (int m = 0; m < 4; ++m)
            for (int c = 0; c < 3; ++c)
                for (int cp = 0; cp < 3; ++cp) {
                    R val = R(0);
                    for (int b = 0; b < 3; ++b)
                        for (int d = 0; d < 3; ++d) val += HF[c + 3 * b][cp + 3 * d] * Q[n][b] * Q[m][d];
                    Hout[3 * n + c][3 * m + cp] = V * val;
                }
}

// ---------------------------------------------------------------------------------------
// Solver state and step
// ---------------------------------------------------------------------------------------

struct Params {
    float dt;
    int32_t numIterations;
    int32_t substeps;
    float gravity[3];
    int32_t useAcceleration;
    float accelerationRho;
    float mu;
    float lambda;
    int32_t initFromInertia;  // 1: x <- y before iterating (VBD_GPU_Wrapper.cu:334)
    int32_t contactOn;
    int32_t contactAxis;
    float contactK;
    float contactPlane;
    int32_t mode;       // 0 closed, 1 faithful block, 2 faithful with reference cost structure
    int32_t threads;    // OpenMP threads inside one color (race-free; result independent of it)
    int32_t energyEvery;  // 10 in the reference; 0 disables the diagnostic
    int32_t material;   // 0 stable Neo-Hookean (Energy.cpp), 1 the GPU wrapper's corotated model
};

template <class R> struct Mesh {
    int nV, nT, nColors;
    const int32_t* tets;
    const int32_t* adjOff;
    const int32_t* adj;
    const int32_t* colorOff;
    const int32_t* colorVerts;
    const uint8_t* fixed;
    const R* mass;
    const R* dminv;  // row-major 9 per tet
    const R* vol;
    const R* tetMu;      // nullable
    const R* tetLambda;  // nullable
};

template <class R> struct Solver {
    std::vector<R> xprev, y, cur, pp;
    std::vector<double> energyHistory;
};

template <class R>
static inline R omega_next(int order, R rho, R prev);

template <>
inline float omega_next<float>(int order, float rho, float prev) {  // VBD.cpp:302-311
    switch (order) {
        case 1: return 1.0f;
        case 2: return (float)(2.0 / (2.0 - (double)(rho * rho)));
        default: return (float)(4.0 / (4.0 - (double)((rho * rho) * prev)));
    }
}
template <>
inline double omega_next<double>(int order, double rho, double prev) {
    switch (order) {
        case 1: return 1.0;
        case 2: return 2.0 / (2.0 - rho * rho);
        default: return 4.0 / (4.0 - rho * rho * prev);
    }
}

template <class R>
static inline R inv_h2(R h);
template <>
inline float inv_h2<float>(float h) { return (float)(1.0 / (double)(h * h)); }  // VBD.cpp:116-117
template <>
inline double inv_h2<double>(double h) { return 1.0 / (h * h); }

template <class R>
static inline R damp_const();
template <> inline float damp_const<float>() { return 0.999f; }   // TetMesh.h:308 (scalar cast to float)
template <> inline double damp_const<double>() { return 0.999; }

// Diagnostic log-Neo-Hookean energy of one tet (NeoHookean.cpp:6-20) with J = prod(sigma)
// = |det F| (singular values are non-negative, NeoHookean.cpp:150-153).
template <class R>
static inline R tet_log_energy(const M3<R>& F, R mu, R lambda) {
    R tr = R(0);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) tr += F(r, c) * F(r, c);
    R J = std::abs(det3(F));
    return (R)(mu * 0.5 * (tr - 3) - mu * std::log(J) + 0.5 * lambda * std::pow(std::log(J), 2.0));
}

template <class R>
static R total_energy(const Mesh<R>& M, const Params& p, const R* x) {  // VBD.cpp:258-288
    R total = R(0);
    for (int t = 0; t < M.nT; ++t) {
        const int32_t* tet = M.tets + 4 * t;
        M3<R> Ds, DmInv, F;
        compute_Dm(Ds, x, tet);
        std::memcpy(&DmInv, M.dminv + 9 * t, sizeof(DmInv));
        matmul(F, Ds, DmInv);
        R mu = M.tetMu ? M.tetMu[t] : (R)p.mu;
        R lam = M.tetLambda ? M.tetLambda[t] : (R)p.lambda;
        total += tet_log_energy(F, mu, lam) * M.vol[t];
    }
    return total;
}

// VBDsolver::solveVertex + the 3x3 solve of VBDsolver::solve for one vertex (VBD.cpp:124-203).
// Returns the position increment applied (zero if the system was singular).
template <class R>
static inline void solve_vertex(const Mesh<R>& M, const Params& p, R* x, const R* y, int v, R s) {
    R H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    R f[3];
    R ms = M.mass[v] * s;
    if (p.material == 1) {
        // VBD_GPU_Wrapper.cu:203-213: coeff = m s; H diagonal coeff * 1.000001, f = coeff (y - x)
        for (int c = 0; c < 3; ++c) H[c][c] += ms * (R)1.000001f;
        for (int c = 0; c < 3; ++c) f[c] = ms * (y[3 * v + c] - x[3 * v + c]);
    } else {
        for (int c = 0; c < 3; ++c) H[c][c] += ms;
        for (int c = 0; c < 3; ++c) f[c] = (M.mass[v] * (y[3 * v + c] - x[3 * v + c])) * s;
    }
    for (int e = M.adjOff[v]; e < M.adjOff[v + 1]; ++e) {
        int t = M.adj[e];
        const int32_t* tet = M.tets + 4 * t;
        int k = -1;
        for (int i = 0; i < 4; ++i)
            if (tet[i] == v) { k = i; break; }
        if (k < 0) continue;
        R mu = M.tetMu ? M.tetMu[t] : (R)p.mu;
        R lam = M.tetLambda ? M.tetLambda[t] : (R)p.lambda;
        M3<R> Ds, DmInv, F;
        compute_Dm(Ds, x, tet);
        std::memcpy(&DmInv, M.dminv + 9 * t, sizeof(DmInv));
        matmul(F, Ds, DmInv);
        R Hl[3][3], gl[3];
        if (p.material == 1) {
            block_corotated(F, DmInv, mu, lam, M.vol[t], k, Hl, gl);
        } else if (p.mode == 0) {
            block_closed(F, DmInv, mu, lam, M.vol[t], k, Hl, gl);
        } else if (p.mode == 1) {
            block_faithful(F, DmInv, mu, lam, M.vol[t], k, Hl, gl);
        } else {
            std::vector<R> material_params = {mu, lam};  // VBD.cpp:220 (allocation kept for cost)
            R Hfull[12][12], gfull[12];
            full_faithful(F, DmInv, material_params[0], material_params[1], M.vol[t], Hfull, gfull);
            for (int r = 0; r < 3; ++r) {
                gl[r] = gfull[3 * k + r];
                for (int c = 0; c < 3; ++c) Hl[r][c] = Hfull[3 * k + r][3 * k + c];
            }
        }
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) H[r][c] += Hl[r][c];
            f[r] -= gl[r];
        }
    }
    if (p.contactOn) {
        int a = p.contactAxis;
        R zp = (R)p.contactPlane, kc = (R)p.contactK;
        R d = zp - x[3 * v + a];
        if (d > R(0)) {
            H[a][a] += kc;
            f[a] += kc * d;
        }
    }
    // VBD.cpp:141-153 then Types.h:85-112
    M3<R> A;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) A(r, c) = H[r][c];
    R detH = det3(A);
    if (!(std::abs((double)detH) > 1e-8)) return;
    R det = detH;  // same formula as AnalyticSolver3x3::solve
    if (std::abs((double)det) < 1e-8) return;
    R invDet = R(1) / det;
    R adj[3][3];
    adj[0][0] = A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1);
    adj[0][1] = A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2);
    adj[0][2] = A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1);
    adj[1][0] = A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2);
    adj[1][1] = A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0);
    adj[1][2] = A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2);
    adj[2][0] = A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0);
    adj[2][1] = A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1);
    adj[2][2] = A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
    for (int r = 0; r < 3; ++r) {
        R dx = invDet * sum3(adj[r][0] * f[0], adj[r][1] * f[1], adj[r][2] * f[2]);
        x[3 * v + r] += dx;
    }
}

// One Gauss-Seidel sweep over all colors (VBD.cpp:114-158, numThreads = 1 semantics: colors
// in ascending order; vertices of one color are independent, so threads inside a color do
// not change the result).
template <class R>
static void sweep(const Mesh<R>& M, const Params& p, R* x, const R* y, R s) {
    for (int c = 0; c < M.nColors; ++c) {
        int b = M.colorOff[c], e = M.colorOff[c + 1];
#pragma omp parallel for num_threads(p.threads > 0 ? p.threads : 1) schedule(dynamic, 64) if (p.threads > 1)
        for (int i = b; i < e; ++i) {
            int v = M.colorVerts[i];
            if (M.fixed[v]) continue;
            solve_vertex(M, p, x, y, v, s);
        }
    }
}

// VBDsolver::step (VBD.cpp:60-101) for one frame.
template <class R>
static void step(const Mesh<R>& M, const Params& p, Solver<R>& S, R* x, R* v) {
    const int N = M.nV;
    if ((int)S.pp.size() != 3 * N) {
        S.xprev.assign(3 * N, R(0));
        S.y.assign(3 * N, R(0));
        S.cur.assign(3 * N, R(0));
        S.pp.assign(3 * N, R(0));  // VBD.cpp:30-31
    }
    R h = (R)p.dt / (R)p.substeps;
    R s = inv_h2<R>(h);
    R g[3] = {(R)p.gravity[0], (R)p.gravity[1], (R)p.gravity[2]};
    for (int sub = 0; sub < p.substeps; ++sub) {
        // forwardStep (VBD.cpp:103-112)
        std::memcpy(S.xprev.data(), x, sizeof(R) * 3 * N);
        for (int i = 0; i < N; ++i)
            if (!M.fixed[i])
                for (int c = 0; c < 3; ++c) v[3 * i + c] += h * g[c];
        for (int i = 0; i < N; ++i)
            for (int c = 0; c < 3; ++c) S.y[3 * i + c] = M.fixed[i] ? x[3 * i + c] : x[3 * i + c] + h * v[3 * i + c];
        if (p.initFromInertia) std::memcpy(x, S.y.data(), sizeof(R) * 3 * N);
        R omega = R(1);
        for (int it = 0; it < p.numIterations; ++it) {
            std::memcpy(S.cur.data(), x, sizeof(R) * 3 * N);
            sweep(M, p, x, S.y.data(), s);
            if (p.useAcceleration) {
                omega = omega_next<R>(it + 1, (R)p.accelerationRho, omega);
                if (omega > 1.0) {
                    for (int i = 0; i < N; ++i)
                        if (!M.fixed[i])
                            for (int c = 0; c < 3; ++c) {
                                R pv = S.pp[3 * i + c];
                                x[3 * i + c] = omega * (x[3 * i + c] - pv) + pv;
                            }
                }
                S.pp.swap(S.cur);
            }
            if (p.mode == 2 && p.energyEvery > 0 && it % p.energyEvery == 0)
                S.energyHistory.push_back((double)total_energy(M, p, x));
        }
        // updateVelocity (TetMesh.h:305-311)
        for (int i = 0; i < N; ++i)
            if (!M.fixed[i])
                for (int c = 0; c < 3; ++c) v[3 * i + c] = (x[3 * i + c] - S.xprev[3 * i + c]) / h * damp_const<R>();
    }
}

template <class R>
static Mesh<R> make_mesh(int nV, int nT, const int32_t* tets, const int32_t* adjOff, const int32_t* adj, int nColors,
                         const int32_t* colorOff, const int32_t* colorVerts, const uint8_t* fixed, const R* mass,
                         const R* dminv, const R* vol, const R* tetMu, const R* tetLambda) {
    Mesh<R> M;
    M.nV = nV; M.nT = nT; M.nColors = nColors;
    M.tets = tets; M.adjOff = adjOff; M.adj = adj;
    M.colorOff = colorOff; M.colorVerts = colorVerts; M.fixed = fixed;
    M.mass = mass; M.dminv = dminv; M.vol = vol; M.tetMu = tetMu; M.tetLambda = tetLambda;
    return M;
}

}  // namespace orc

using namespace orc;

extern "C" {

int orc_params_size() { return (int)sizeof(Params); }

// TetMesh::initialize (TetMesh.h:213-275): rest volume, lumped mass, vertex->tet adjacency in
// ascending tet order.  adj_off has nV+1 entries, adj has 4*nT entries.
void orc_mesh_init_f32(int nV, const float* rest, int nT, const int32_t* tets, float density, float* mass,
                       float* tetRestVol, int32_t* adjOff, int32_t* adj) {
    std::vector<int32_t> deg(nV, 0);
    for (int i = 0; i < nV; ++i) mass[i] = 0.0f;
    for (int t = 0; t < nT; ++t) {
        const int32_t* v = tets + 4 * t;
        float e0[3], e1[3], e2[3];
        for (int c = 0; c < 3; ++c) {
            e0[c] = rest[3 * v[1] + c] - rest[3 * v[0] + c];
            e1[c] = rest[3 * v[2] + c] - rest[3 * v[0] + c];
            e2[c] = rest[3 * v[3] + c] - rest[3 * v[0] + c];
        }
        // e0.dot(e1.cross(e2)) (TetMesh.h:297)
        float cx = e1[1] * e2[2] - e1[2] * e2[1];
        float cy = e1[2] * e2[0] - e1[0] * e2[2];
        float cz = e1[0] * e2[1] - e1[1] * e2[0];
        float vol = std::abs(sum3(e0[0] * cx, e0[1] * cy, e0[2] * cz)) / 6.0f;
        tetRestVol[t] = vol;
        float tetMass = vol * density;
        for (int j = 0; j < 4; ++j) {
            deg[v[j]]++;
            mass[v[j]] += tetMass / 4.0f;
        }
    }
    adjOff[0] = 0;
    for (int i = 0; i < nV; ++i) adjOff[i + 1] = adjOff[i] + deg[i];
    std::vector<int32_t> cursor(adjOff, adjOff + nV);
    for (int t = 0; t < nT; ++t)
        for (int j = 0; j < 4; ++j) adj[cursor[tets[4 * t + j]]++] = t;
}

void orc_mesh_init_f64(int nV, const float* rest, int nT, const int32_t* tets, double density, double* mass,
                       double* tetRestVol) {
    for (int i = 0; i < nV; ++i) mass[i] = 0.0;
    for (int t = 0; t < nT; ++t) {
        const int32_t* v = tets + 4 * t;
        double e0[3], e1[3], e2[3];
        for (int c = 0; c < 3; ++c) {
            e0[c] = (double)rest[3 * v[1] + c] - rest[3 * v[0] + c];
            e1[c] = (double)rest[3 * v[2] + c] - rest[3 * v[0] + c];
            e2[c] = (double)rest[3 * v[3] + c] - rest[3 * v[0] + c];
        }
        double cx = e1[1] * e2[2] - e1[2] * e2[1];
        double cy = e1[2] * e2[0] - e1[0] * e2[2];
        double cz = e1[0] * e2[1] - e1[1] * e2[0];
        double vol = std::abs(e0[0] * cx + (e0[1] * cy + e0[2] * cz)) / 6.0;
        tetRestVol[t] = vol;
        for (int j = 0; j < 4; ++j) mass[v[j]] += vol * density / 4.0;
    }
}

// GraphColoring::colorVertices (GraphColoring.cpp:28-66): greedy first-fit in ascending
// vertex order; a vertex's neighbours are the other corners of its incident tets.
// Returns the number of colors; color_off must hold nV+1 entries, color_verts nV entries.
int orc_color(int nV, const int32_t* tets, const int32_t* adjOff, const int32_t* adj, int32_t* colorOf,
              int32_t* colorOff, int32_t* colorVerts) {
    for (int i = 0; i < nV; ++i) colorOf[i] = -1;
    std::vector<std::vector<int32_t>> groups;
    std::vector<char> used;
    for (int v = 0; v < nV; ++v) {
        used.assign(groups.size() + 2, 0);
        for (int e = adjOff[v]; e < adjOff[v + 1]; ++e) {
            const int32_t* tet = tets + 4 * adj[e];
            for (int i = 0; i < 4; ++i) {
                int u = tet[i];
                if (u != v && colorOf[u] != -1 && colorOf[u] < (int)used.size()) used[colorOf[u]] = 1;
            }
        }
        int c = 0;
        while (c < (int)used.size() && used[c]) ++c;
        colorOf[v] = c;
        if (c >= (int)groups.size()) groups.resize(c + 1);
        groups[c].push_back(v);
    }
    int k = 0;
    colorOff[0] = 0;
    for (size_t c = 0; c < groups.size(); ++c) {
        for (int v : groups[c]) colorVerts[k++] = v;
        colorOff[c + 1] = k;
    }
    return (int)groups.size();
}

// VBDsolver::precomputeTetData (VBD.cpp:36-58): DmInv (row-major) and V0 per tet.
void orc_precompute_f32(const float* rest, int nT, const int32_t* tets, float* dminv, float* vol) {
    for (int t = 0; t < nT; ++t) {
        M3<float> Dm;
        compute_Dm(Dm, rest, tets + 4 * t);
        M3<float> inv = inverse3(Dm);
        std::memcpy(dminv + 9 * t, &inv, sizeof(inv));
        // compute_Volume (Energy.cpp:24-32): |d3 . (d1 x d2)| / 6
        float d1[3] = {Dm(0, 0), Dm(1, 0), Dm(2, 0)}, d2[3] = {Dm(0, 1), Dm(1, 1), Dm(2, 1)},
              d3[3] = {Dm(0, 2), Dm(1, 2), Dm(2, 2)};
        float cx = d1[1] * d2[2] - d1[2] * d2[1];
        float cy = d1[2] * d2[0] - d1[0] * d2[2];
        float cz = d1[0] * d2[1] - d1[1] * d2[0];
        vol[t] = std::abs(sum3(d3[0] * cx, d3[1] * cy, d3[2] * cz)) / 6.0f;
    }
}

void orc_precompute_f64(const float* rest, int nT, const int32_t* tets, double* dminv, double* vol) {
    for (int t = 0; t < nT; ++t) {
        M3<double> Dm;
        const int32_t* tet = tets + 4 * t;
        for (int j = 0; j < 3; ++j)
            for (int c = 0; c < 3; ++c) Dm(c, j) = (double)rest[3 * tet[j + 1] + c] - (double)rest[3 * tet[0] + c];
        M3<double> inv = inverse3(Dm);
        std::memcpy(dminv + 9 * t, &inv, sizeof(inv));
        vol[t] = std::abs(det3(Dm)) / 6.0;
    }
}

// Per-tet element quantities for tests: F from the 4 corner positions, then either the full
// 12-gradient / 12x12 Hessian (Energy.cpp:90-158) or the closed-form (k,k) block.
void orc_tet_full_f64(const double* x12, const double* dminv9, double mu, double lambda, double V, double* H144,
                      double* g12) {
    int32_t tet[4] = {0, 1, 2, 3};
    M3<double> Ds, DmInv, F;
    compute_Dm(Ds, x12, tet);
    std::memcpy(&DmInv, dminv9, sizeof(DmInv));
    matmul(F, Ds, DmInv);
    double H[12][12], g[12];
    full_faithful(F, DmInv, mu, lambda, V, H, g);
    std::memcpy(H144, H, sizeof(H));
    std::memcpy(g12, g, sizeof(g));
}

void orc_tet_full_f32(const float* x12, const float* dminv9, float mu, float lambda, float V, float* H144, float* g12) {
    int32_t tet[4] = {0, 1, 2, 3};
    M3<float> Ds, DmInv, F;
    compute_Dm(Ds, x12, tet);
    std::memcpy(&DmInv, dminv9, sizeof(DmInv));
    matmul(F, Ds, DmInv);
    float H[12][12], g[12];
    full_faithful(F, DmInv, mu, lambda, V, H, g);
    std::memcpy(H144, H, sizeof(H));
    std::memcpy(g12, g, sizeof(g));
}

void orc_tet_block_f64(const double* x12, const double* dminv9, double mu, double lambda, double V, int k,
                       int faithful, double* H9, double* g3) {
    int32_t tet[4] = {0, 1, 2, 3};
    M3<double> Ds, DmInv, F;
    compute_Dm(Ds, x12, tet);
    std::memcpy(&DmInv, dminv9, sizeof(DmInv));
    matmul(F, Ds, DmInv);
    double H[3][3], g[3];
    if (faithful) block_faithful(F, DmInv, mu, lambda, V, k, H, g);
    else block_closed(F, DmInv, mu, lambda, V, k, H, g);
    std::memcpy(H9, H, sizeof(H));
    std::memcpy(g3, g, sizeof(g));
}

// Corotated block (material 1) and the polar factor it uses, in double.
void orc_tet_block_corot_f64(const double* x12, const double* dminv9, double mu, double lambda, double V, int k,
                             double* H9, double* g3) {
    int32_t tet[4] = {0, 1, 2, 3};
    M3<double> Ds, DmInv, F;
    compute_Dm(Ds, x12, tet);
    std::memcpy(&DmInv, dminv9, sizeof(DmInv));
    matmul(F, Ds, DmInv);
    double H[3][3], g[3];
    block_corotated(F, DmInv, mu, lambda, V, k, H, g);
    std::memcpy(H9, H, sizeof(H));
    std::memcpy(g3, g, sizeof(g));
}

void orc_polar_rotation_f64(const double* F9, double* R9) {
    M3<double> F;
    std::memcpy(&F, F9, sizeof(F));
    M3<double> Rm = polar_rotation(F);
    std::memcpy(R9, &Rm, sizeof(Rm));
}

void orc_tet_block_f32(const float* x12, const float* dminv9, float mu, float lambda, float V, int k, int faithful,
                       float* H9, float* g3) {
    int32_t tet[4] = {0, 1, 2, 3};
    M3<float> Ds, DmInv, F;
    compute_Dm(Ds, x12, tet);
    std::memcpy(&DmInv, dminv9, sizeof(DmInv));
    matmul(F, Ds, DmInv);
    float H[3][3], g[3];
    if (faithful) block_faithful(F, DmInv, mu, lambda, V, k, H, g);
    else block_closed(F, DmInv, mu, lambda, V, k, H, g);
    std::memcpy(H9, H, sizeof(H));
    std::memcpy(g3, g, sizeof(g));
}

// Stable Neo-Hookean energy density (the potential whose gradient is Energy.cpp:97):
// Psi = mu/2 (|F|^2 - 3) - mu (J - 1) + lambda/2 (J - 1)^2, times V.  Used by FD tests.
double orc_tet_energy_f64(const double* x12, const double* dminv9, double mu, double lambda, double V) {
    int32_t tet[4] = {0, 1, 2, 3};
    M3<double> Ds, DmInv, F;
    compute_Dm(Ds, x12, tet);
    std::memcpy(&DmInv, dminv9, sizeof(DmInv));
    matmul(F, Ds, DmInv);
    double tr = 0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) tr += F(r, c) * F(r, c);
    double J = det3(F);
    return V * (0.5 * mu * (tr - 3.0) - mu * (J - 1.0) + 0.5 * lambda * (J - 1.0) * (J - 1.0));
}

double orc_omega_f32(int order, float rho, float prev) { return (double)omega_next<float>(order, rho, prev); }

// Run nFrames calls of VBDsolver::step.  x, v (3N, in/out).  frames_out (nullable) receives
// the positions after every frame (nFrames x 3N).  Returns elapsed wall seconds.
double orc_simulate_f32(const Params* p, int nV, int nT, const int32_t* tets, const int32_t* adjOff,
                        const int32_t* adj, int nColors, const int32_t* colorOff, const int32_t* colorVerts,
                        const uint8_t* fixed, const float* mass, const float* dminv, const float* vol,
                        const float* tetMu, const float* tetLambda, float* x, float* v, int nFrames,
                        float* framesOut) {
    Mesh<float> M = make_mesh<float>(nV, nT, tets, adjOff, adj, nColors, colorOff, colorVerts, fixed, mass, dminv,
                                     vol, tetMu, tetLambda);
    Solver<float> S;
    auto t0 = std::chrono::steady_clock::now();
    for (int f = 0; f < nFrames; ++f) {
        step(M, *p, S, x, v);
        if (framesOut) std::memcpy(framesOut + (size_t)f * 3 * nV, x, sizeof(float) * 3 * nV);
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

double orc_simulate_f64(const Params* p, int nV, int nT, const int32_t* tets, const int32_t* adjOff,
                        const int32_t* adj, int nColors, const int32_t* colorOff, const int32_t* colorVerts,
                        const uint8_t* fixed, const double* mass, const double* dminv, const double* vol,
                        const double* tetMu, const double* tetLambda, double* x, double* v, int nFrames,
                        double* framesOut) {
    Mesh<double> M = make_mesh<double>(nV, nT, tets, adjOff, adj, nColors, colorOff, colorVerts, fixed, mass,
                                       dminv, vol, tetMu, tetLambda);
    Solver<double> S;
    auto t0 = std::chrono::steady_clock::now();
    for (int f = 0; f < nFrames; ++f) {
        step(M, *p, S, x, v);
        if (framesOut) std::memcpy(framesOut + (size_t)f * 3 * nV, x, sizeof(double) * 3 * nV);
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Single color pass restricted to a vertex subset: for each listed vertex (assumed to be of
// one color, so no listed vertex is a neighbour of another), solve its local system against
// positions x and write the new position to xout (3 per listed vertex).  x is not modified.
// Used to check one GPU color launch on meshes too large for a whole CPU run.
double orc_solve_subset_f32(const Params* p, int nV, int nT, const int32_t* tets, const int32_t* adjOff,
                            const int32_t* adj, const uint8_t* fixed, const float* mass, const float* dminv,
                            const float* vol, const float* tetMu, const float* tetLambda, const float* x,
                            const float* y, int nSub, const int32_t* verts, float* xout) {
    Mesh<float> M = make_mesh<float>(nV, nT, tets, adjOff, adj, 0, nullptr, nullptr, fixed, mass, dminv, vol, tetMu,
                                     tetLambda);
    float h = p->dt / (float)p->substeps;
    float s = inv_h2<float>(h);
    std::vector<float> xl(x, x + 3 * (size_t)nV);
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for num_threads(p->threads > 0 ? p->threads : 1) schedule(dynamic, 16) if (p->threads > 1)
    for (int i = 0; i < nSub; ++i) {
        int v = verts[i];
        if (!fixed[v]) solve_vertex(M, *p, xl.data(), y, v, s);
    }
    double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < nSub; ++i)
        for (int c = 0; c < 3; ++c) xout[3 * i + c] = xl[3 * (size_t)verts[i] + c];
    return el;
}

// One sweep in the reference's own scheme (VBD.cpp:119-121): OpenMP over the colors, each color
// swept by one thread while other colors run concurrently.  Vertices of different colors that
// share a tet race, so the result is not deterministic; for timing only.  Returns seconds.
double orc_sweep_over_colors_f32(const Params* p, int nV, int nT, const int32_t* tets, const int32_t* adjOff,
                                 const int32_t* adj, int nColors, const int32_t* colorOff, const int32_t* colorVerts,
                                 const uint8_t* fixed, const float* mass, const float* dminv, const float* vol,
                                 const float* tetMu, const float* tetLambda, float* x, const float* y) {
    Mesh<float> M = make_mesh<float>(nV, nT, tets, adjOff, adj, nColors, colorOff, colorVerts, fixed, mass, dminv, vol,
                                     tetMu, tetLambda);
    float h = p->dt / (float)p->substeps;
    float s = inv_h2<float>(h);
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for num_threads(p->threads > 0 ? p->threads : 1) schedule(dynamic)
    for (int c = 0; c < nColors; ++c)
        for (int i = colorOff[c]; i < colorOff[c + 1]; ++i) {
            int v = colorVerts[i];
            if (!fixed[v]) solve_vertex(M, *p, x, y, v, s);
        }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Total diagnostic energy (VBD.cpp:258-288).
double orc_total_energy_f32(const Params* p, int nT, const int32_t* tets, const float* dminv, const float* vol,
                            const float* tetMu, const float* tetLambda, const float* x) {
    Mesh<float> M = make_mesh<float>(0, nT, tets, nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr, dminv, vol,
                                     tetMu, tetLambda);
    return (double)total_energy(M, *p, x);
}

}  // extern "C"
