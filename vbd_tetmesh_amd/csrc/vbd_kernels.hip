// gfx950 device code of the VBD solver: mesh build (CSR, color lists, color-major
// renumbering, pair records), the per-color vertex solve, Chebyshev blend, time-step
// kernels and the diagnostic energy reduction.
//
// Algorithm references (extinguish7/VBD_TetMesh):
//   per-vertex solve ............ VBD.cpp:160-203 (inertia + incident tets), :141-153 (3x3 solve)
//   per-(tet, corner) block ..... Energy.cpp:90-158 in closed form (see block comment below)
//   3x3 solve ................... Types.h:85-112
//   Chebyshev blend ............. VBD.cpp:290-300
//   predict / velocity .......... TetMesh.h:305-348
//   adjacency / colors .......... TetMesh.h:254-266, GraphColoring.cpp:28-66, VBD_GPU_Wrapper.cu:286-308
#include "vbd_internal.h"

#include <algorithm>
#include <cmath>

namespace vbd {

// ---------------------------------------------------------------------------------------
// Wave-level sums with DPP.  gsum<L> sums each aligned group of L lanes of a wave64; the
// total is valid in the group's last lane (lane % L == L-1).  row_shr steps build suffix
// sums inside a 16-lane row; row_bcast15 / row_bcast31 carry row totals across rows.
// Lanes whose DPP source is outside the row (or whose row is masked off) receive 0.
// ---------------------------------------------------------------------------------------
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_take(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xF, false));
}

template <int L>
__device__ __forceinline__ float gsum(float v) {
    if constexpr (L >= 2) v += dpp_take<0x111, 0xF>(v);  // row_shr:1
    if constexpr (L >= 4) v += dpp_take<0x112, 0xF>(v);  // row_shr:2
    if constexpr (L >= 8) v += dpp_take<0x114, 0xF>(v);  // row_shr:4
    if constexpr (L >= 16) v += dpp_take<0x118, 0xF>(v); // row_shr:8
    if constexpr (L >= 32) v += dpp_take<0x142, 0xA>(v); // row_bcast:15 into rows 1, 3
    if constexpr (L >= 64) v += dpp_take<0x143, 0xC>(v); // row_bcast:31 into rows 2, 3
    return v;
}

// Blocks of a launch are dealt round-robin to the 8 XCDs (block b runs on XCD b % 8).  The
// remap hands each XCD one contiguous run of logical blocks, so neighbouring vertices (and
// the neighbour positions they gather) share one XCD's L2.  gridDim.x is a multiple of 8.
__device__ __forceinline__ int xcd_block() {
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// ---------------------------------------------------------------------------------------
// Per-(vertex, tet) contribution.
//
// F is formed as the reference forms it, F = Ds*DmInv with Ds = [x1-x0, x2-x0, x3-x0]
// (Energy.cpp:11-21), rewritten without a corner index: a pair record holds three corner
// slots (a, b, c), V0, and three DmInv rows (R0, R1, R2) such that
//   vertex is corner 0 (flag bit set):  a, b, c = corners 1, 2, 3;  R = rows 0, 1, 2;
//                                       F = sum_j (y_j - x_s) R_j^T,  q = -(R0+R1+R2)
//   vertex is corner k > 0:             a = corner 0, b, c = the remaining corners in tet
//                                       order; R0 = row k-1, R1, R2 = their rows;
//                                       F = (x_s - y_a) R0^T + (y_b - y_a) R1^T + (y_c - y_a) R2^T,
//                                       q = R0
// where q is the vertex's shape vector, row k of Q = Ps*DmInv (Energy.cpp:101-106).  Then
// (Energy.cpp:93-97)  J = det F, C = cof F, P = mu(F - C) + lambda(J-1) C,
//   g = V P q,  H = V [mu |q|^2 I + lambda (C q)(C q)^T]
// and H is exactly the vertex's 3x3 diagonal block of the 12x12 Hessian of
// Energy.cpp:115-154: the (lambda(J-1) - mu) d^2J/dF^2 term contracts with q q^T to zero.
// ---------------------------------------------------------------------------------------
// Record slot fields hold 27-bit slots.  Bit 31 of .x flags "vertex is corner 0"; bits 27-30
// of .x, .y, .z together hold the record's owner as an offset (12 bits) from the owner of the
// first record of its 4096-record chunk (chunkBase); the tile build reads it and rewrites it,
// in the padded tile records, as the owner's index inside its tile.
constexpr int SLOT_MASK = 0x07FFFFFF;
constexpr int CHUNK_SHIFT = 12;

__device__ __forceinline__ int owner_offset(const i4 t) {
    return ((t.x >> 27) & 15) | (((t.y >> 27) & 15) << 4) | (((t.z >> 27) & 15) << 8);
}

struct Acc {
    float h00, h01, h02, h11, h12, h22, r0, r1, r2;
};

struct Kin {  // F, its cofactor C and J of one tet, seen from one of its corners
    float F00, F01, F02, F10, F11, F12, F20, F21, F22;
    float C00, C01, C02, C10, C11, C12, C20, C21, C22;
    float J, q0, q1, q2;
};

__device__ __forceinline__ Kin pair_kinematics(const bool corner0, const f4 a, const f4 b, const float c,
                                               const f4 xs, const f4 ya, const f4 yb, const f4 yc) {
    // R0 = (a.x, a.y, a.z), R1 = (a.w, b.x, b.y), R2 = (b.z, b.w, c)
    Kin k;
    const f4 base = corner0 ? xs : ya;
    const f4 first = corner0 ? ya : xs;
    const float e0x = first.x - base.x, e0y = first.y - base.y, e0z = first.z - base.z;
    const float e1x = yb.x - base.x, e1y = yb.y - base.y, e1z = yb.z - base.z;
    const float e2x = yc.x - base.x, e2y = yc.y - base.y, e2z = yc.z - base.z;
    k.F00 = e0x * a.x + e1x * a.w + e2x * b.z;
    k.F01 = e0x * a.y + e1x * b.x + e2x * b.w;
    k.F02 = e0x * a.z + e1x * b.y + e2x * c;
    k.F10 = e0y * a.x + e1y * a.w + e2y * b.z;
    k.F11 = e0y * a.y + e1y * b.x + e2y * b.w;
    k.F12 = e0y * a.z + e1y * b.y + e2y * c;
    k.F20 = e0z * a.x + e1z * a.w + e2z * b.z;
    k.F21 = e0z * a.y + e1z * b.x + e2z * b.w;
    k.F22 = e0z * a.z + e1z * b.y + e2z * c;
    // cofactor (Energy.cpp:61-73) and J
    k.C00 = k.F11 * k.F22 - k.F12 * k.F21;
    k.C01 = k.F12 * k.F20 - k.F10 * k.F22;
    k.C02 = k.F10 * k.F21 - k.F11 * k.F20;
    k.C10 = k.F21 * k.F02 - k.F22 * k.F01;
    k.C11 = k.F22 * k.F00 - k.F20 * k.F02;
    k.C12 = k.F20 * k.F01 - k.F21 * k.F00;
    k.C20 = k.F01 * k.F12 - k.F02 * k.F11;
    k.C21 = k.F02 * k.F10 - k.F00 * k.F12;
    k.C22 = k.F00 * k.F11 - k.F01 * k.F10;
    k.J = k.F00 * k.C00 + k.F01 * k.C01 + k.F02 * k.C02;
    k.q0 = corner0 ? -(a.x + a.w + b.z) : a.x;
    k.q1 = corner0 ? -(a.y + b.x + b.w) : a.y;
    k.q2 = corner0 ? -(a.z + b.y + c) : a.z;
    return k;
}

__device__ __forceinline__ void pair_accumulate(const bool corner0, const f4 a, const f4 b, const float c,
                                                const float V, const float tmu, const float tlam, const f4 xs,
                                                const f4 ya, const f4 yb, const f4 yc, Acc& A) {
    const Kin k = pair_kinematics(corner0, a, b, c, xs, ya, yb, yc);
    const float aj = tlam * (k.J - 1.0f);
    const float q0 = k.q0, q1 = k.q1, q2 = k.q2;
    // P q = mu (F q - C q) + aj C q
    const float Cq0 = k.C00 * q0 + k.C01 * q1 + k.C02 * q2;
    const float Cq1 = k.C10 * q0 + k.C11 * q1 + k.C12 * q2;
    const float Cq2 = k.C20 * q0 + k.C21 * q1 + k.C22 * q2;
    const float Fq0 = k.F00 * q0 + k.F01 * q1 + k.F02 * q2;
    const float Fq1 = k.F10 * q0 + k.F11 * q1 + k.F12 * q2;
    const float Fq2 = k.F20 * q0 + k.F21 * q1 + k.F22 * q2;
    const float g0 = tmu * (Fq0 - Cq0) + aj * Cq0;
    const float g1 = tmu * (Fq1 - Cq1) + aj * Cq1;
    const float g2 = tmu * (Fq2 - Cq2) + aj * Cq2;
    const float qq = q0 * q0 + q1 * q1 + q2 * q2;
    const float vm = V * tmu * qq;
    const float vl = V * tlam;
    const float vc0 = vl * Cq0, vc1 = vl * Cq1, vc2 = vl * Cq2;
    A.h00 += vm + vc0 * Cq0;
    A.h01 += vc0 * Cq1;
    A.h02 += vc0 * Cq2;
    A.h11 += vm + vc1 * Cq1;
    A.h12 += vc1 * Cq2;
    A.h22 += vm + vc2 * Cq2;
    A.r0 -= V * g0;
    A.r1 -= V * g1;
    A.r2 -= V * g2;
}

// Neighbour position gather through a raw buffer resource: 32-bit byte offsets instead of
// 64-bit address arithmetic per load.
__device__ __forceinline__ f4 gather_pos(__amdgpu_buffer_rsrc_t rsrc, int slot) {
    return __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(rsrc, slot << 4, 0, 0));
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t pos_rsrc(const f4* pos, int n) {
    // stride 0 (raw), num_records in bytes, flags = default data format for gfx9 raw loads
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<f4*>(pos), (short)0, n * 16, 0x00020000);
}

// ---------------------------------------------------------------------------------------
// Per-color vertex solve.
//
// One group of LPV lanes per free vertex of the color; lane j of the group takes the
// vertex's records j, j+LPV, ...; each pass issues the loads of U records (and then their
// neighbour gathers) before any arithmetic so more memory requests are in flight.  The group
// sums the 6 distinct H entries and 3 force entries (DPP), the last lane adds the inertia
// term m/h^2 (VBD.cpp:170-173), optional penalty contact, and applies the adjugate 3x3 solve
// when |det H| > 1e-8 (VBD.cpp:141-153, Types.h:85-112).  With SAVE the vertex's position
// before its update (its position at the start of the sweep) is written to `save` for the
// Chebyshev blend.
// ---------------------------------------------------------------------------------------
template <int LPV, int U, bool MAT, bool CONTACT, bool SAVE>
__global__ __launch_bounds__(256) void k_color_solve(int slot0, int count, int nslots, f4* __restrict__ pos,
                                                     const f4* __restrict__ inert, const float* __restrict__ mass,
                                                     const int* __restrict__ row, const i4* __restrict__ ptet,
                                                     const f4* __restrict__ pa, const f4* __restrict__ pb,
                                                     const float* __restrict__ pc, const f2* __restrict__ pmat,
                                                     f4* __restrict__ save, float mu, float lambda, float sdt,
                                                     ContactArgs ct) {
    const int gpb = blockDim.x / LPV;
    const int g = xcd_block() * gpb + (int)threadIdx.x / LPV;
    const int lane = (int)threadIdx.x % LPV;
    const bool valid = g < count;
    const int s = slot0 + g;

    Acc A = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    f4 xs = {0.f, 0.f, 0.f, 0.f};
    if (valid) {
        const __amdgpu_buffer_rsrc_t prs = pos_rsrc(pos, nslots);
        xs = pos[s];
        const int p0 = row[s], p1 = row[s + 1];
        for (int base = p0 + lane; base < p1; base += LPV * U) {
            i4 t[U];
            f4 a[U], b[U];
            float c[U];
            f2 mm[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int p = base + u * LPV;
                if (p < p1) {
                    t[u] = __builtin_nontemporal_load(&ptet[p]);
                    a[u] = __builtin_nontemporal_load(&pa[p]);
                    b[u] = __builtin_nontemporal_load(&pb[p]);
                    c[u] = __builtin_nontemporal_load(&pc[p]);
                    if constexpr (MAT) mm[u] = __builtin_nontemporal_load(&pmat[p]);
                }
            }
            f4 ya[U], yb[U], yc[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (base + u * LPV < p1) {
                    ya[u] = gather_pos(prs, t[u].x & SLOT_MASK);
                    yb[u] = gather_pos(prs, t[u].y & SLOT_MASK);
                    yc[u] = gather_pos(prs, t[u].z & SLOT_MASK);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (base + u * LPV < p1) {
                    float tmu = mu, tlam = lambda;
                    if constexpr (MAT) {
                        tmu = mm[u].x;
                        tlam = mm[u].y;
                    }
                    pair_accumulate(t[u].x < 0, a[u], b[u], c[u], __int_as_float(t[u].w), tmu, tlam, xs, ya[u], yb[u],
                                    yc[u], A);
                }
            }
        }
    }
    float h00 = gsum<LPV>(A.h00), h01 = gsum<LPV>(A.h01), h02 = gsum<LPV>(A.h02);
    float h11 = gsum<LPV>(A.h11), h12 = gsum<LPV>(A.h12), h22 = gsum<LPV>(A.h22);
    float r0 = gsum<LPV>(A.r0), r1 = gsum<LPV>(A.r1), r2 = gsum<LPV>(A.r2);
    if (!valid || lane != LPV - 1) return;

    const float m = mass[s];
    const f4 ys = inert[s];
    const float ms = m * sdt;
    h00 += ms; h11 += ms; h22 += ms;
    r0 += (m * (ys.x - xs.x)) * sdt;
    r1 += (m * (ys.y - xs.y)) * sdt;
    r2 += (m * (ys.z - xs.z)) * sdt;
    if constexpr (CONTACT) {
        const float xa = ct.axis == 0 ? xs.x : (ct.axis == 1 ? xs.y : xs.z);
        const float d = ct.plane - xa;
        if (d > 0.f) {
            if (ct.axis == 0) { h00 += ct.k; r0 += ct.k * d; }
            else if (ct.axis == 1) { h11 += ct.k; r1 += ct.k * d; }
            else { h22 += ct.k; r2 += ct.k * d; }
        }
    }
    const float det = h00 * (h11 * h22 - h12 * h12) - h01 * (h01 * h22 - h12 * h02) + h02 * (h01 * h12 - h11 * h02);
    f4 xn = xs;
    if (fabsf(det) > 1e-8f) {
        const float inv = 1.0f / det;
        const float a00 = h11 * h22 - h12 * h12, a01 = h02 * h12 - h01 * h22, a02 = h01 * h12 - h02 * h11;
        const float a11 = h00 * h22 - h02 * h02, a12 = h02 * h01 - h00 * h12, a22 = h00 * h11 - h01 * h01;
        xn.x = xs.x + inv * (a00 * r0 + a01 * r1 + a02 * r2);
        xn.y = xs.y + inv * (a01 * r0 + a11 * r1 + a12 * r2);
        xn.z = xs.z + inv * (a02 * r0 + a12 * r1 + a22 * r2);
    }
    if constexpr (SAVE) save[s] = xs;
    pos[s] = xn;
}

// Vertex update from the summed elastic Hessian block h and force r: inertia term m/h^2
// (VBD.cpp:170-173), optional penalty contact, adjugate 3x3 solve when |det H| > 1e-8
// (VBD.cpp:141-153, Types.h:85-112).  Returns the new position.
template <bool CONTACT>
__device__ __forceinline__ f4 vertex_solve(const f4 xs, const f4 ys, const float m, float h00, float h01, float h02,
                                           float h11, float h12, float h22, float r0, float r1, float r2, float sdt,
                                           const ContactArgs& ct) {
#pragma clang fp contract(off)
    const float ms = m * sdt;
    h00 += ms; h11 += ms; h22 += ms;
    r0 += (m * (ys.x - xs.x)) * sdt;
    r1 += (m * (ys.y - xs.y)) * sdt;
    r2 += (m * (ys.z - xs.z)) * sdt;
    if constexpr (CONTACT) {
        const float xa = ct.axis == 0 ? xs.x : (ct.axis == 1 ? xs.y : xs.z);
        const float d = ct.plane - xa;
        if (d > 0.f) {
            if (ct.axis == 0) { h00 += ct.k; r0 += ct.k * d; }
            else if (ct.axis == 1) { h11 += ct.k; r1 += ct.k * d; }
            else { h22 += ct.k; r2 += ct.k * d; }
        }
    }
    const float a00 = __builtin_fmaf(h11, h22, -(h12 * h12));
    const float a01 = __builtin_fmaf(h02, h12, -(h01 * h22));
    const float a02 = __builtin_fmaf(h01, h12, -(h02 * h11));
    const float a11 = __builtin_fmaf(h00, h22, -(h02 * h02));
    const float a12 = __builtin_fmaf(h02, h01, -(h00 * h12));
    const float a22 = __builtin_fmaf(h00, h11, -(h01 * h01));
    const float det = __builtin_fmaf(h00, a00, __builtin_fmaf(h01, a01, h02 * a02));
    f4 xn = xs;
    if (fabsf(det) > 1e-8f) {
        const float inv = 1.0f / det;
        xn.x = __builtin_fmaf(inv, __builtin_fmaf(a00, r0, __builtin_fmaf(a01, r1, a02 * r2)), xs.x);
        xn.y = __builtin_fmaf(inv, __builtin_fmaf(a01, r0, __builtin_fmaf(a11, r1, a12 * r2)), xs.y);
        xn.z = __builtin_fmaf(inv, __builtin_fmaf(a02, r0, __builtin_fmaf(a12, r1, a22 * r2)), xs.z);
    }
    return xn;
}

// ---------------------------------------------------------------------------------------
// Corotated linear elasticity of the reference GPU kernel (VBD_GPU_Wrapper.cu:127-233 with
// VBD_DeviceMath.cuh:127-198), on the full records:
//   R = 5 steps of R <- (R + cof(R) / det R) / 2 from R = F (stop if |det R| < 1e-8),
//   P = 2 mu (F - R) + lambda tr(R^T F - I) R,   g = V P q,
//   H = V [2 mu |q|^2 I + lambda (R q)(R q)^T]   (dR/dF dropped),
// and the wrapper's vertex update: H diagonal m/h^2 * 1.000001, force (m/h^2)(y - x)
// (VBD_GPU_Wrapper.cu:203-213), adjugate solve skipped when |det H| < 1e-8.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void corot_accumulate(const bool corner0, const f4 a, const f4 b, const float c,
                                                 const float V, const float tmu, const float tlam, const f4 xs,
                                                 const f4 ya, const f4 yb, const f4 yc, Acc& A) {
    // Products are fused explicitly (fmaf) and the reciprocal of det is the hardware's v_rcp_f32
    // (1 ulp): about a third fewer VALU instructions than separate multiplies and adds with an
    // IEEE division, and every instantiation still rounds alike (no contraction by the compiler).
#pragma clang fp contract(off)
    const Kin k = pair_kinematics(corner0, a, b, c, xs, ya, yb, yc);
    // Higham steps R <- (R + R^-T) / 2 carried as U_k = 2^k R_k, so that a step is
    // U <- U + cof(U) * (4^k / det U) without halving every entry: cof(U) = 4^k cof(R),
    // det U = 8^k det R, and the singularity test scales by 8^k; all the scalings are powers of
    // two (exact), and R = 2^-k U at the end.  The first step reuses cof(F).
    float R00 = k.F00, R01 = k.F01, R02 = k.F02, R10 = k.F10, R11 = k.F11, R12 = k.F12;
    float R20 = k.F20, R21 = k.F21, R22 = k.F22;
    float C00 = k.C00, C01 = k.C01, C02 = k.C02, C10 = k.C10, C11 = k.C11, C12 = k.C12;
    float C20 = k.C20, C21 = k.C21, C22 = k.C22;
    float back = 1.0f;
#pragma unroll
    for (int it = 0; it < 5; ++it) {
        if (it > 0) {
            C00 = __builtin_fmaf(R11, R22, -(R12 * R21));
            C01 = __builtin_fmaf(R12, R20, -(R10 * R22));
            C02 = __builtin_fmaf(R10, R21, -(R11 * R20));
            C10 = __builtin_fmaf(R21, R02, -(R22 * R01));
            C11 = __builtin_fmaf(R22, R00, -(R20 * R02));
            C12 = __builtin_fmaf(R20, R01, -(R21 * R00));
            C20 = __builtin_fmaf(R01, R12, -(R02 * R11));
            C21 = __builtin_fmaf(R02, R10, -(R00 * R12));
            C22 = __builtin_fmaf(R00, R11, -(R01 * R10));
        }
        // det U by its first row and cofactors (U^-T = cof U / det U)
        const float det = __builtin_fmaf(R00, C00, __builtin_fmaf(R01, C01, R02 * C02));
        if (fabsf(det) < 1e-8f * (float)(1 << (3 * it))) break;
        const float inv = (float)(1 << (2 * it)) * __builtin_amdgcn_rcpf(det);
        R00 = __builtin_fmaf(C00, inv, R00); R01 = __builtin_fmaf(C01, inv, R01);
        R02 = __builtin_fmaf(C02, inv, R02); R10 = __builtin_fmaf(C10, inv, R10);
        R11 = __builtin_fmaf(C11, inv, R11); R12 = __builtin_fmaf(C12, inv, R12);
        R20 = __builtin_fmaf(C20, inv, R20); R21 = __builtin_fmaf(C21, inv, R21);
        R22 = __builtin_fmaf(C22, inv, R22);
        back *= 0.5f;
    }
    R00 *= back; R01 *= back; R02 *= back;
    R10 *= back; R11 *= back; R12 *= back;
    R20 *= back; R21 *= back; R22 *= back;
    const float tr = __builtin_fmaf(R00, k.F00, __builtin_fmaf(R01, k.F01, __builtin_fmaf(R02, k.F02,
                     __builtin_fmaf(R10, k.F10, __builtin_fmaf(R11, k.F11, __builtin_fmaf(R12, k.F12,
                     __builtin_fmaf(R20, k.F20, __builtin_fmaf(R21, k.F21, R22 * k.F22)))))))) - 3.0f;
    const float m2 = 2.0f * tmu, lt = tlam * tr;
    const float q0 = k.q0, q1 = k.q1, q2 = k.q2;
    const float P00 = __builtin_fmaf(m2, k.F00 - R00, lt * R00), P01 = __builtin_fmaf(m2, k.F01 - R01, lt * R01);
    const float P02 = __builtin_fmaf(m2, k.F02 - R02, lt * R02), P10 = __builtin_fmaf(m2, k.F10 - R10, lt * R10);
    const float P11 = __builtin_fmaf(m2, k.F11 - R11, lt * R11), P12 = __builtin_fmaf(m2, k.F12 - R12, lt * R12);
    const float P20 = __builtin_fmaf(m2, k.F20 - R20, lt * R20), P21 = __builtin_fmaf(m2, k.F21 - R21, lt * R21);
    const float P22 = __builtin_fmaf(m2, k.F22 - R22, lt * R22);
    const float rq0 = __builtin_fmaf(R00, q0, __builtin_fmaf(R01, q1, R02 * q2));
    const float rq1 = __builtin_fmaf(R10, q0, __builtin_fmaf(R11, q1, R12 * q2));
    const float rq2 = __builtin_fmaf(R20, q0, __builtin_fmaf(R21, q1, R22 * q2));
    const float d = m2 * __builtin_fmaf(q0, q0, __builtin_fmaf(q1, q1, q2 * q2));
    const float l0 = tlam * rq0, l1 = tlam * rq1, l2 = tlam * rq2;
    A.h00 = __builtin_fmaf(V, __builtin_fmaf(l0, rq0, d), A.h00);
    A.h01 = __builtin_fmaf(V, l0 * rq1, A.h01);
    A.h02 = __builtin_fmaf(V, l0 * rq2, A.h02);
    A.h11 = __builtin_fmaf(V, __builtin_fmaf(l1, rq1, d), A.h11);
    A.h12 = __builtin_fmaf(V, l1 * rq2, A.h12);
    A.h22 = __builtin_fmaf(V, __builtin_fmaf(l2, rq2, d), A.h22);
    A.r0 = __builtin_fmaf(-V, __builtin_fmaf(P00, q0, __builtin_fmaf(P01, q1, P02 * q2)), A.r0);
    A.r1 = __builtin_fmaf(-V, __builtin_fmaf(P10, q0, __builtin_fmaf(P11, q1, P12 * q2)), A.r1);
    A.r2 = __builtin_fmaf(-V, __builtin_fmaf(P20, q0, __builtin_fmaf(P21, q1, P22 * q2)), A.r2);
}

// The wrapper's vertex update (VBD_GPU_Wrapper.cu:203-213): H diagonal + m/h^2 * 1.000001,
// force + (m/h^2)(y - x), penalty contact, adjugate solve unless |det H| < 1e-8.
template <bool CONTACT>
__device__ __forceinline__ f4 corot_vertex_solve(const f4 xs, const f4 ys, const float m, float h00, float h01,
                                                 float h02, float h11, float h12, float h22, float r0, float r1,
                                                 float r2, float sdt, const ContactArgs& ct) {
#pragma clang fp contract(off)
    const float coeff = m * sdt;
    const float reg = coeff * 1.000001f;
    h00 += reg; h11 += reg; h22 += reg;
    r0 += coeff * (ys.x - xs.x);
    r1 += coeff * (ys.y - xs.y);
    r2 += coeff * (ys.z - xs.z);
    if constexpr (CONTACT) {
        const float xa = ct.axis == 0 ? xs.x : (ct.axis == 1 ? xs.y : xs.z);
        const float dd = ct.plane - xa;
        if (dd > 0.f) {
            if (ct.axis == 0) { h00 += ct.k; r0 += ct.k * dd; }
            else if (ct.axis == 1) { h11 += ct.k; r1 += ct.k * dd; }
            else { h22 += ct.k; r2 += ct.k * dd; }
        }
    }
    const float det = h00 * (h11 * h22 - h12 * h12) - h01 * (h01 * h22 - h12 * h02) + h02 * (h01 * h12 - h11 * h02);
    f4 xn = xs;
    if (!(fabsf(det) < 1e-8f)) {
        const float inv = 1.0f / det;
        const float a00 = h11 * h22 - h12 * h12, a01 = h02 * h12 - h01 * h22, a02 = h01 * h12 - h02 * h11;
        const float a11 = h00 * h22 - h02 * h02, a12 = h02 * h01 - h00 * h12, a22 = h00 * h11 - h01 * h01;
        xn.x = xs.x + inv * (a00 * r0 + a01 * r1 + a02 * r2);
        xn.y = xs.y + inv * (a01 * r0 + a11 * r1 + a12 * r2);
        xn.z = xs.z + inv * (a02 * r0 + a12 * r1 + a22 * r2);
    }
    return xn;
}

// ---------------------------------------------------------------------------------------
// Tiled per-color vertex solve on compact records.
//
// Per (vertex s, tet) record the closed-form block needs only F q, C q and J (q = the
// vertex's shape vector, row s of Q = Ps DmInv, Energy.cpp:101-106).  With a, b, c the other
// three corners (a = corner 0, or corner 1 when s is corner 0), o = corner 0 of the tet
// (x_s or y_a), delta = x_s - y_a, u = y_b - y_a, w = y_c - y_a:
//   F q = Ds (DmInv q) = beta delta + alpha1 (y_b - o) + alpha2 (y_c - o)
//         (the edges of Ds from corner 0 as in Energy.cpp:11-21; beta = +-alpha0; the alpha
//          are precomputed)
//   C q = dJ/dx_s = d (u x w),   d = det(DmInv) * (-1, +1, -1, +1)[corner of s]
//   J   = (C q) . delta                                         (J is affine in x_s and
//                                                                vanishes on the face a b c)
// and then, exactly as in the full form,
//   g = V [mu (F q - C q) + lambda (J-1) C q],   H = mu V |q|^2 I + lambda V (C q)(C q)^T,
// with |q|^2 = alpha0 = beta (s not corner 0) or -(alpha0 + alpha1 + alpha2) = beta - alpha1
// - alpha2 (s corner 0, q = -(R0 + R1 + R2)).  The record is ptet (3 slots, V) + pq (beta,
// alpha1, alpha2, d): 32 B instead of 52 B, and about half the arithmetic of forming F and
// cof(F).
//
// The color's free vertices are cut into tiles of consecutive slots holding at most
// RT = K * blockDim.x records and fewer than blockDim.x / 8 vertices (tileStart, built on the
// host).  Each tile's records are copied into RT padded slots at tile * RT (tpt, tpq, ...;
// padding records are all zero), so a block finds its records from its index alone.  One
// block per tile; XCD x takes the x-th eighth of the tiles.
//   1. thread i takes records i, i+T, ... (K of them) of its tile: the whole block streams one
//      contiguous range, every lane busy.  The record's owner is tileStart[tile] + the
//      tile-local vertex index held in the owner bits.  The 9 contribution values (6 Hessian
//      entries, 3 force entries) go to LDS, 9 consecutive words per record (an odd stride, so
//      consecutive records fall in distinct banks); the tile's CSR row offsets, positions,
//      inertia targets and masses are staged in LDS too.
//   2. after a barrier, TG = 4 lanes per vertex sum its records (lane l: records l, l+TG, ...
//      of the vertex, then a DPP tree), so the summation order depends only on the vertex's
//      valence, not on the tile cut, the grid or the block size: results are bitwise
//      identical across (K, blockDim) choices.  The group's last lane solves and writes the
//      vertex.
// ---------------------------------------------------------------------------------------
constexpr int TG = 4;    // lanes per vertex in the reduction stage
constexpr int VDIV = 8;  // a tile holds fewer than blockDim.x / VDIV vertices
constexpr int BUF_NT = 2;  // nontemporal cache policy bit of the raw buffer loads

template <int K, bool MAT>
struct RecRegs {
    i4 t[K];
    f4 q[K];     // compact coefficients (stable Neo-Hookean)
    f4 a[K], b[K];  // full DmInv record (corotated): a, b, c as in the per-vertex kernels
    float c[K];
    f2 mm[K];
};

struct RecSrc {  // buffer resources over one color's records (byte offsets from its first record)
    __amdgpu_buffer_rsrc_t t, q, a, b, c, mm;
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rec_rsrc(const void* base, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}

// Loads are unconditional (index clamped to the color's last record): no branch around
// them, so the compiler's wait counts can keep the gathers in flight with the vertex loads.
template <int K, bool MAT, bool COROT>
__device__ __forceinline__ void load_records(RecRegs<K, MAT>& R, const RecSrc& src, int r0, int pmax, int tid, int T) {
    // every record's corner slots first: the gathers need only those, and vmcnt counts loads in
    // issue order, so a coefficient load issued between two slot loads would hold back the
    // second record's gathers
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int p = min(r0 + tid + k * T, pmax);  // relative to the color's first record
        R.t[k] = __builtin_bit_cast(i4, __builtin_amdgcn_raw_buffer_load_b128(src.t, p << 4, 0, BUF_NT));
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int p = min(r0 + tid + k * T, pmax);
        if constexpr (COROT) {
            R.a[k] = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(src.a, p << 4, 0, BUF_NT));
            R.b[k] = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(src.b, p << 4, 0, BUF_NT));
            R.c[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(src.c, p << 2, 0, BUF_NT));
        } else {
            R.q[k] = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(src.q, p << 4, 0, BUF_NT));
        }
        if constexpr (MAT)
            R.mm[k] = __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(src.mm, p << 3, 0, BUF_NT));
    }
}

// Workgroup barrier ordering LDS only.  __syncthreads() also orders global memory, which on
// this target waits for every outstanding vector load and store.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Explicit fmaf and no contraction: every instantiation (K, MAT, ...) rounds alike, so the
// tiled configurations give bitwise identical results.
__device__ __forceinline__ Acc compact_contrib(const bool corner0, const f4 qv, const float V, const float tmu,
                                               const float tlam, const f4 xs, const f4 ya, const f4 yb,
                                               const f4 yc) {
#pragma clang fp contract(off)
    const float qq = corner0 ? (qv.x - qv.y) - qv.z : qv.x;  // |q|^2
    const float dx = xs.x - ya.x, dy = xs.y - ya.y, dz = xs.z - ya.z;
    const float ux = yb.x - ya.x, uy = yb.y - ya.y, uz = yb.z - ya.z;
    const float wx = yc.x - ya.x, wy = yc.y - ya.y, wz = yc.z - ya.z;
    const f4 o = corner0 ? xs : ya;
    const float bx = yb.x - o.x, by = yb.y - o.y, bz = yb.z - o.z;
    const float cx = yc.x - o.x, cy = yc.y - o.y, cz = yc.z - o.z;
    const float c0 = qv.w * __builtin_fmaf(uy, wz, -(uz * wy));
    const float c1 = qv.w * __builtin_fmaf(uz, wx, -(ux * wz));
    const float c2 = qv.w * __builtin_fmaf(ux, wy, -(uy * wx));
    const float J = __builtin_fmaf(c0, dx, __builtin_fmaf(c1, dy, c2 * dz));
    const float f0 = __builtin_fmaf(qv.x, dx, __builtin_fmaf(qv.y, bx, qv.z * cx));
    const float f1 = __builtin_fmaf(qv.x, dy, __builtin_fmaf(qv.y, by, qv.z * cy));
    const float f2v = __builtin_fmaf(qv.x, dz, __builtin_fmaf(qv.y, bz, qv.z * cz));
    // g/V = mu (F q - C q) + lambda (J - 1) C q; F q - C q first (near-exact cancellation)
    const float aj = tlam * (J - 1.0f);
    const float g0 = __builtin_fmaf(tmu, f0 - c0, aj * c0);
    const float g1 = __builtin_fmaf(tmu, f1 - c1, aj * c1);
    const float g2 = __builtin_fmaf(tmu, f2v - c2, aj * c2);
    const float vm = tmu * (V * qq);
    const float vl = V * tlam;
    const float v0 = vl * c0, v1 = vl * c1, v2 = vl * c2;
    Acc A;
    A.h00 = __builtin_fmaf(v0, c0, vm);
    A.h01 = v0 * c1;
    A.h02 = v0 * c2;
    A.h11 = __builtin_fmaf(v1, c1, vm);
    A.h12 = v1 * c2;
    A.h22 = __builtin_fmaf(v2, c2, vm);
    A.r0 = -(V * g0);
    A.r1 = -(V * g1);
    A.r2 = -(V * g2);
    return A;
}

// One tile of one color: records, gathers, contributions, the LDS reduction and the vertex
// solves (see the block comment above).  `tileStart` points at the color's tile table and the
// record pointers at the color's first padded record slot; T is the block size.  Uses every
// thread of the block; the waves past the tile's last vertex group skip the reduction (no early
// return, so a persistent caller can loop over tiles).
template <int K, bool MAT, bool CONTACT, bool SAVE, bool COROT>
__device__ __forceinline__ void tile_solve(const int tile, const int T, const int* __restrict__ tileStart,
                                           const int nslots,
                                           f4* __restrict__ pos, const f4* __restrict__ inert,
                                           const float* __restrict__ mass, const int* __restrict__ row,
                                           const i4* __restrict__ ptet, const f4* __restrict__ pq,
                                           const f4* __restrict__ pa, const f4* __restrict__ pb,
                                           const float* __restrict__ pc, const f2* __restrict__ pmat,
                                           f4* __restrict__ save, const float mu, const float lambda, const float sdt,
                                           const ContactArgs& ct, float* red) {
    // LDS: 9 * RT contribution values, then VM = T/VDIV entries (a tile has at most VM - 1
    // vertices): row offsets (nv + 1 of them), x, y, m
    const int RT = K * T;
    const int VM = T / VDIV;
    int* lrow = reinterpret_cast<int*>(red + 9 * RT);
    f4* lx = reinterpret_cast<f4*>(red + 9 * RT + VM);
    f4* ly = lx + VM;
    float* lm = reinterpret_cast<float*>(ly + VM);
    const int tid = (int)threadIdx.x;
    const __amdgpu_buffer_rsrc_t prs = pos_rsrc(pos, nslots);
    // the tile's RT padded record slots: their address depends on the tile index alone, so the
    // record loads issue first, with nothing to wait for
    const size_t base = (size_t)tile * RT;
    RecSrc src;
    src.t = rec_rsrc(ptet + base, 16 * (int64_t)RT);
    if constexpr (COROT) {
        src.a = rec_rsrc(pa + base, 16 * (int64_t)RT);
        src.b = rec_rsrc(pb + base, 16 * (int64_t)RT);
        src.c = rec_rsrc(pc + base, 4 * (int64_t)RT);
    } else {
        src.q = rec_rsrc(pq + base, 16 * (int64_t)RT);
    }
    src.mm = MAT ? rec_rsrc(pmat + base, 8 * (int64_t)RT) : src.t;
    RecRegs<K, MAT> R;
    load_records<K, MAT, COROT>(R, src, 0, RT - 1, tid, T);
    const int v0 = tileStart[tile], v1 = tileStart[tile + 1];
    __builtin_amdgcn_sched_barrier(0);  // the tile-table load issues with the record loads
    const int nv = v1 - v0;
    // the corners first: they need only the record's slots, while the record's own position
    // needs the tile's first slot from the scalar tile-table load
    f4 xs[K], ya[K], yb[K], yc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        ya[k] = gather_pos(prs, R.t[k].x & SLOT_MASK);
        yb[k] = gather_pos(prs, R.t[k].y & SLOT_MASK);
        yc[k] = gather_pos(prs, R.t[k].z & SLOT_MASK);
    }
    __builtin_amdgcn_sched_barrier(0);  // keeps the corner gathers ahead of the tile-table wait
#pragma unroll
    for (int k = 0; k < K; ++k) xs[k] = gather_pos(prs, v0 + owner_offset(R.t[k]));  // owner: tile-local index
    // the tile's vertex data, loaded by the first wave only (a tile has fewer than 64 vertices);
    // clamped indices, lanes past nv load and ignore a copy.  Issued after the gathers, so that the
    // first wave's gathers are in flight with them (vmcnt counts loads in issue order: loads issued
    // before the gathers would hold the first wave's gathers back until they returned).
    const bool wave0 = __builtin_amdgcn_readfirstlane(tid >> 6) == 0;
    int vrow = 0;
    f4 vx = {}, vy = {};
    float vm = 0.f;
    if (wave0) {
        vrow = row[min(v0 + tid, v1)];
        const int vs = min(v0 + tid, v1 - 1);
        vx = pos[vs];
        vy = __builtin_nontemporal_load(&inert[vs]);
        vm = __builtin_nontemporal_load(&mass[vs]);
    }
    const int vrow0 = row[v0];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = tid + k * T;
        float tmu = mu, tlam = lambda;
        if constexpr (MAT) {
            tmu = R.mm[k].x;
            tlam = R.mm[k].y;
        }
        Acc A;
        if constexpr (COROT) {
            A = Acc{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            corot_accumulate(R.t[k].x < 0, R.a[k], R.b[k], R.c[k], __int_as_float(R.t[k].w), tmu, tlam, xs[k], ya[k],
                             yb[k], yc[k], A);
        } else {
            A = compact_contrib(R.t[k].x < 0, R.q[k], __int_as_float(R.t[k].w), tmu, tlam, xs[k], ya[k], yb[k], yc[k]);
        }
        {  // padding records contribute zeros past the last vertex's range
            float* rq = red + 9 * q;  // record-major: one address, immediate offsets
            rq[0] = A.h00;
            rq[1] = A.h01;
            rq[2] = A.h02;
            rq[3] = A.h11;
            rq[4] = A.h12;
            rq[5] = A.h22;
            rq[6] = A.r0;
            rq[7] = A.r1;
            rq[8] = A.r2;
        }
    }
    if (tid <= nv) lrow[tid] = vrow - vrow0;
    if (tid < nv) {
        lx[tid] = vx;
        ly[tid] = vy;
        lm[tid] = vm;
    }
    lds_barrier();

    // one group of TG lanes per tile vertex; waves past the tile's last group have nothing to
    // sum (the test is uniform per wave)
    if ((tid & ~63) < nv * TG) {
        const int lane = tid % TG;
        const int i = tid / TG;
        const bool ok = i < nv;
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f, s5 = 0.f, s6 = 0.f, s7 = 0.f, s8 = 0.f;
        if (ok) {
            auto add9 = [&](const float* rq) __attribute__((always_inline)) {
                s0 += rq[0];
                s1 += rq[1];
                s2 += rq[2];
                s3 += rq[3];
                s4 += rq[4];
                s5 += rq[5];
                s6 += rq[6];
                s7 += rq[7];
                s8 += rq[8];
            };
            // the lane's records q, q + TG, ... added in that order; two per step, so that the
            // LDS reads of both are in flight together (the same sums, half the round trips)
            const int q1 = lrow[i + 1];
            int q = lrow[i] + lane;
            for (; q + TG < q1; q += 2 * TG) {
                add9(red + 9 * q);
                add9(red + 9 * (q + TG));
            }
            if (q < q1) add9(red + 9 * q);
        }
        s0 = gsum<TG>(s0); s1 = gsum<TG>(s1); s2 = gsum<TG>(s2);
        s3 = gsum<TG>(s3); s4 = gsum<TG>(s4); s5 = gsum<TG>(s5);
        s6 = gsum<TG>(s6); s7 = gsum<TG>(s7); s8 = gsum<TG>(s8);
        if (ok && lane == TG - 1) {
            const f4 x = lx[i];
            f4 xn;
            if constexpr (COROT)
                xn = corot_vertex_solve<CONTACT>(x, ly[i], lm[i], s0, s1, s2, s3, s4, s5, s6, s7, s8, sdt, ct);
            else
                xn = vertex_solve<CONTACT>(x, ly[i], lm[i], s0, s1, s2, s3, s4, s5, s6, s7, s8, sdt, ct);
            if constexpr (SAVE) save[v0 + i] = x;
            pos[v0 + i] = xn;
        }
    }
}

// COROT: the corotated material of the reference GPU kernel on the full records (pa, pb, pc)
// instead of the compact stable Neo-Hookean records (pq); same tiles, reduction and order.
template <int K, bool MAT, bool CONTACT, bool SAVE, bool COROT>
// The arguments a block needs before its first loads come first (the record streams, the tile
// table, the positions, the tile and slot counts and the block size T, which the kernel takes as
// an argument instead of reading the dispatch's workgroup size): built with kernel-argument
// preloading they arrive in SGPRs, so the record loads issue without a scalar load before them.
__global__ __launch_bounds__(512) void k_color_tiled(const i4* __restrict__ ptet, const f4* __restrict__ pq,
                                                     const f2* __restrict__ pmat, const int* __restrict__ tileStart,
                                                     f4* __restrict__ pos, int ntiles, int nslots, int T,
                                                     const f4* __restrict__ inert, const float* __restrict__ mass,
                                                     const int* __restrict__ row, const f4* __restrict__ pa,
                                                     const f4* __restrict__ pb, const float* __restrict__ pc,
                                                     f4* __restrict__ save, float mu, float lambda, float sdt,
                                                     ContactArgs ct) {
    extern __shared__ float red[];
    const int xcd = (int)(blockIdx.x & 7);
    const int tile = (int)(((int64_t)ntiles * xcd) >> 3) + (int)(blockIdx.x >> 3);
    if (tile >= (int)(((int64_t)ntiles * (xcd + 1)) >> 3)) return;  // whole block
    tile_solve<K, MAT, CONTACT, SAVE, COROT>(tile, T, tileStart, nslots, pos, inert, mass, row, ptet, pq, pa, pb, pc,
                                             pmat, save, mu, lambda, sdt, ct, red);
}

template <int K, bool MAT, bool CONTACT, bool SAVE, bool COROT>
static hipError_t launch_tiled_t(const DevMesh& m, int color, int block, float sdt, float mu, float lambda,
                                 const ContactArgs& ct, f4* save, hipStream_t s) {
    const int ntiles = m.colorTileCount[color];
    if (ntiles <= 0) return hipSuccess;
    const size_t lds = sizeof(float) * 9 * (size_t)K * block +
                       (sizeof(int) + 2 * sizeof(f4) + sizeof(float)) * (size_t)(block / VDIV);
    const int per = (ntiles + 7) / 8;
    if (COROT && !m.tpa) return hipErrorInvalidValue;  // full records not padded (build_tiles full)
    const size_t base = (size_t)m.colorTileBase[color] * m.tileRecords;
    hipLaunchKernelGGL((k_color_tiled<K, MAT, CONTACT, SAVE, COROT>), dim3(8 * per), dim3(block), lds, s,
                       m.tpt + base, m.tpq + base, MAT ? m.tpm + base : nullptr, m.tileStart + m.colorTileBase[color],
                       m.pos, ntiles, m.nV, block, m.inert, m.mass, m.row, COROT ? m.tpa + base : nullptr,
                       COROT ? m.tpb + base : nullptr, COROT ? m.tpc + base : nullptr, save, mu, lambda, sdt, ct);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Persistent sweep: one launch runs every color of one Gauss-Seidel sweep (and the Chebyshev
// blend after it) with a grid barrier between the phases, instead of one launch per color.
// Block b takes the tiles L, L + G, ... of each color (L: the XCD-aware logical index, G the
// grid), through the same tile_solve as k_color_tiled, so the results are bitwise those of the
// per-color launches.
//
// Grid barrier, hierarchical over the 8 XCDs (each with its own L2): blocks b with the same
// b & 7 form a group (the hardware dispatches block b to XCD b mod 8).  Words of `bar`, each on
// its own 128-byte line (PERSIST_LINE words apart): 0 the group arrivals, 1 the generation,
// 2 the timeout flag, 3 + x the arrivals of group x.
//  * every wave waits for its own stores (vmcnt 0): they are in its XCD's L2;
//  * the block's first thread reads the generation and arrives at its group; a block running on
//    another XCD than its group's (read from the XCC_ID register) writes its own L2 back first;
//  * the last arrival of a group resets the group count, writes its XCD's L2 back (agent-scope
//    release: one write-back per XCD instead of one per wave) and arrives globally; the last of
//    the 8 resets the global count and bumps the generation; every other block spins on the
//    generation with s_sleep, at most `spin` polls;
//  * then the first thread acquires (its CU's L1 and its XCD's L2 are invalidated) and waits for
//    the invalidation before the block barrier.
// The counts are back to zero after every barrier, so graph replays need no reset.  A spin that
// runs out raises the flag and the block returns; every barrier after it then times out too and
// the launch ends; the host checks the flag.
// ---------------------------------------------------------------------------------------
#define VBD_WAIT_VMEM() __builtin_amdgcn_s_waitcnt(0x0F70)  // s_waitcnt vmcnt(0) (gfx9 encoding)

__device__ __forceinline__ bool grid_barrier(unsigned* bar, unsigned nblocks, unsigned spin) {
    __shared__ unsigned ok;
    VBD_WAIT_VMEM();  // this wave's stores have reached the L2
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned* const cnt = bar;
        unsigned* const gen = bar + PERSIST_LINE;
        unsigned* const flag = bar + 2 * PERSIST_LINE;
        const unsigned x = blockIdx.x & 7;
        unsigned* const xcnt = bar + (3 + x) * PERSIST_LINE;
        const unsigned g0 = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        VBD_WAIT_VMEM();  // the generation is read before this block arrives
        const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 0xf;  // HW_REG_XCC_ID
        if (xcc != x) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // not in the group's L2
        bool waiter = true;
        if (__hip_atomic_fetch_add(xcnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == (nblocks >> 3) - 1) {
            __hip_atomic_store(xcnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // the group's positions, device-wide
            if (__hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 7u) {
                __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                waiter = false;
            }
        }
        unsigned good = 1;
        if (waiter) {
            unsigned n = 0;
            while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g0) {
                if (++n > spin || __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    __hip_atomic_store(flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    good = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the other XCDs' positions
        VBD_WAIT_VMEM();  // the invalidation is done before the block's other waves load
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

template <int K, bool MAT, bool CONTACT, bool SAVE>
__global__ __launch_bounds__(512) void k_sweep_persistent(int nColors, const int* __restrict__ colorBase,
                                                          const int* __restrict__ colorTiles,
                                                          const int* __restrict__ tileStart, int nslots, int nFree,
                                                          f4* __restrict__ pos, const f4* __restrict__ inert,
                                                          const float* __restrict__ mass, const int* __restrict__ row,
                                                          const i4* __restrict__ ptet, const f4* __restrict__ pq,
                                                          const f2* __restrict__ pmat, f4* __restrict__ save,
                                                          const f4* __restrict__ pp, float omega, float mu,
                                                          float lambda, float sdt, ContactArgs ct,
                                                          unsigned* __restrict__ bar, unsigned spin) {
    extern __shared__ __attribute__((aligned(16))) float red[];  // after the barrier's word: f4 arrays inside
    const unsigned G = gridDim.x;
    const int per = (int)(G >> 3);
    const int xcd = (int)(blockIdx.x & 7), j = (int)(blockIdx.x >> 3);
    const int RT = K * (int)blockDim.x;
    for (int c = 0; c < nColors; ++c) {
        const int n = colorTiles[c];
        const int cb = colorBase[c];
        // XCD x takes the x-th eighth of the color's tiles, its blocks strided through them
        const int lo = (int)(((int64_t)n * xcd) >> 3), hi = (int)(((int64_t)n * (xcd + 1)) >> 3);
        bool first = true;
        for (int t = lo + j; t < hi; t += per) {
            if (!first) lds_barrier();  // the previous tile's LDS reads are done
            first = false;
            const size_t rb = (size_t)cb * RT;
            tile_solve<K, MAT, CONTACT, SAVE, false>(t, (int)blockDim.x, tileStart + cb, nslots, pos, inert, mass, row,
                                                     ptet + rb,
                                                     pq + rb, nullptr, nullptr, nullptr, MAT ? pmat + rb : nullptr,
                                                     save, mu, lambda, sdt, ct, red);
        }
        if (c + 1 < nColors || omega > 1.0f)
            if (!grid_barrier(bar, G, spin)) return;
    }
    if (omega > 1.0f) {  // the Chebyshev blend (VBD.cpp:290-300), rounded as k_blend rounds it
#pragma clang fp contract(off)
        for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < nFree; i += (int)(G * blockDim.x)) {
            const f4 x = pos[i];
            const f4 q = pp[i];
            f4 r = x;
            r.x = omega * (x.x - q.x) + q.x;
            r.y = omega * (x.y - q.y) + q.y;
            r.z = omega * (x.z - q.z) + q.z;
            pos[i] = r;
        }
    }
}

template <int K, bool MAT, bool CONTACT, bool SAVE>
static hipError_t persistent_t(const DevMesh& m, int block, int grid, float sdt, float mu, float lambda,
                               const ContactArgs& ct, f4* save, const f4* pp, float omega, unsigned* bar,
                               unsigned spin, hipStream_t s) {
    const size_t lds = sizeof(float) * 9 * (size_t)K * block +
                       (sizeof(int) + 2 * sizeof(f4) + sizeof(float)) * (size_t)(block / VDIV);
    hipLaunchKernelGGL((k_sweep_persistent<K, MAT, CONTACT, SAVE>), dim3(grid), dim3(block), lds, s, m.nColors,
                       m.dColorTileBase, m.dColorTileCount, m.tileStart, m.nV, m.nFree, m.pos, m.inert, m.mass, m.row,
                       m.tpt, m.tpq, m.tpm, save, pp, omega, mu, lambda, sdt, ct, bar, spin);
    return hipGetLastError();
}

// Resident blocks of the persistent sweep per CU (occupancy) for (K, block), 0 if unsupported.
template <int K, bool MAT, bool CONTACT, bool SAVE>
static int persistent_occupancy_t(int block) {
    const size_t lds = sizeof(float) * 9 * (size_t)K * block +
                       (sizeof(int) + 2 * sizeof(f4) + sizeof(float)) * (size_t)(block / VDIV);
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(&k_sweep_persistent<K, MAT, CONTACT, SAVE>),
                                                     block, lds) != hipSuccess)
        return 0;
    return n;
}

#define VBD_PERSIST_DISPATCH(FN, ...)                                                             \
    do {                                                                                         \
        const bool mat_ = m.perTetMat, con_ = ct.on != 0, sv_ = save != nullptr;                 \
        if (per_thread == 1) {                                                                   \
            if (!mat_ && !con_ && !sv_) return FN<1, false, false, false>(__VA_ARGS__);          \
            if (!mat_ && !con_ && sv_) return FN<1, false, false, true>(__VA_ARGS__);            \
            if (!mat_ && con_ && !sv_) return FN<1, false, true, false>(__VA_ARGS__);            \
            if (!mat_ && con_ && sv_) return FN<1, false, true, true>(__VA_ARGS__);              \
            if (mat_ && !con_ && !sv_) return FN<1, true, false, false>(__VA_ARGS__);            \
            if (mat_ && !con_ && sv_) return FN<1, true, false, true>(__VA_ARGS__);              \
            if (mat_ && con_ && !sv_) return FN<1, true, true, false>(__VA_ARGS__);              \
            return FN<1, true, true, true>(__VA_ARGS__);                                         \
        }                                                                                        \
        if (!mat_ && !con_ && !sv_) return FN<2, false, false, false>(__VA_ARGS__);              \
        if (!mat_ && !con_ && sv_) return FN<2, false, false, true>(__VA_ARGS__);                \
        if (!mat_ && con_ && !sv_) return FN<2, false, true, false>(__VA_ARGS__);                \
        if (!mat_ && con_ && sv_) return FN<2, false, true, true>(__VA_ARGS__);                  \
        if (mat_ && !con_ && !sv_) return FN<2, true, false, false>(__VA_ARGS__);                \
        if (mat_ && !con_ && sv_) return FN<2, true, false, true>(__VA_ARGS__);                  \
        if (mat_ && con_ && !sv_) return FN<2, true, true, false>(__VA_ARGS__);                  \
        return FN<2, true, true, true>(__VA_ARGS__);                                             \
    } while (0)

int persistent_occupancy(const DevMesh& m, int per_thread, int block, const ContactArgs& ct, bool save_on) {
    if (!tiled_config_supported(per_thread, block)) return 0;
    const f4* save = save_on ? reinterpret_cast<const f4*>(1) : nullptr;
    VBD_PERSIST_DISPATCH(persistent_occupancy_t, block);
}

hipError_t launch_sweep_persistent(const DevMesh& m, int per_thread, int block, int grid, float sdt, float mu,
                                   float lambda, const ContactArgs& ct, f4* save, const f4* pp, float omega,
                                   unsigned* bar, unsigned spin, hipStream_t s) {
    if (!tiled_config_supported(per_thread, block) || m.tileRecords != per_thread * block ||
        m.tileVerts != block / VDIV - 1 || !m.tileStart || !m.tpt || !m.dColorTileBase || grid < 8 || (grid & 7) ||
        !bar || (omega > 1.0f && !pp))
        return hipErrorInvalidValue;
    VBD_PERSIST_DISPATCH(persistent_t, m, block, grid, sdt, mu, lambda, ct, save, pp, omega, bar, spin, s);
}
#undef VBD_PERSIST_DISPATCH

template <int K, bool COROT>
static hipError_t launch_tiled_k(const DevMesh& m, int color, int block, float sdt, float mu, float lambda,
                                 const ContactArgs& ct, f4* save, hipStream_t s) {
    const bool mat = m.perTetMat, con = ct.on != 0, sv = save != nullptr;
#define VBD_DISPATCH(MT, CT, SV)                                                                                   \
    if (mat == MT && con == CT && sv == SV)                                                                        \
        return launch_tiled_t<K, MT, CT, SV, COROT>(m, color, block, sdt, mu, lambda, ct, save, s);
    VBD_DISPATCH(false, false, false) VBD_DISPATCH(false, false, true)
    VBD_DISPATCH(false, true, false) VBD_DISPATCH(false, true, true)
    VBD_DISPATCH(true, false, false) VBD_DISPATCH(true, false, true)
    VBD_DISPATCH(true, true, false) VBD_DISPATCH(true, true, true)
#undef VBD_DISPATCH
    return hipErrorInvalidValue;
}

bool tiled_config_supported(int per_thread, int block) {
    if (block < 64 || block > 512 || (block % 64) != 0) return false;
    return per_thread == 1 || per_thread == 2;
}

hipError_t launch_color_tiled(const DevMesh& m, int color, int per_thread, int block, float sdt, float mu,
                              float lambda, const ContactArgs& ct, f4* save, hipStream_t s, bool corotated) {
    if (!tiled_config_supported(per_thread, block) || m.tileRecords != per_thread * block ||
        m.tileVerts != block / VDIV - 1 || !m.tileStart || !m.tpt)
        return hipErrorInvalidValue;
    if (corotated)
        return per_thread == 1 ? launch_tiled_k<1, true>(m, color, block, sdt, mu, lambda, ct, save, s)
                               : launch_tiled_k<2, true>(m, color, block, sdt, mu, lambda, ct, save, s);
    return per_thread == 1 ? launch_tiled_k<1, false>(m, color, block, sdt, mu, lambda, ct, save, s)
                           : launch_tiled_k<2, false>(m, color, block, sdt, mu, lambda, ct, save, s);
}

// Greedy cut of every color's slot range into tiles of at most `records` records and
// `max_verts` vertices (host; uses the host copy of the CSR row offsets).  Fails if a vertex
// alone exceeds the record cap.
// Padded tile records: slot j of tile entry i (an index into tileStart / tileRec, a color's
// closing entry included) holds record tileRec[i] + j of the CSR order when j is below the
// tile's record count, with the owner bits rewritten to the tile-local vertex index; other
// slots are zero (a zero-volume record: no contribution, and no vertex sums it).
__global__ __launch_bounds__(256) void k_pad_records(int nentries, int rt, const int* __restrict__ tileStart,
                                                     const int* __restrict__ tileRec,
                                                     const int* __restrict__ chunkBase, const i4* __restrict__ ptet,
                                                     const f4* __restrict__ pq, const f2* __restrict__ pmat,
                                                     const f4* __restrict__ pa, const f4* __restrict__ pb,
                                                     const float* __restrict__ pc, i4* __restrict__ tpt,
                                                     f4* __restrict__ tpq, f2* __restrict__ tpm,
                                                     f4* __restrict__ tpa, f4* __restrict__ tpb,
                                                     float* __restrict__ tpc) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)nentries * rt) return;
    const int i = (int)(g / rt), j = (int)(g % rt);
    const int r0 = tileRec[i];
    const int cnt = i + 1 < nentries ? tileRec[i + 1] - r0 : 0;
    i4 t = i4{0, 0, 0, 0};
    f4 q = f4{0.f, 0.f, 0.f, 0.f}, a = q, b = q;
    f2 mm = f2{0.f, 0.f};
    float c = 0.f;
    if (j < cnt) {
        const int p = r0 + j;
        t = ptet[p];
        const int local = chunkBase[p >> CHUNK_SHIFT] + owner_offset(t) - tileStart[i];
        t.x = (t.x & ~(15 << 27)) | ((local & 15) << 27);
        t.y = (t.y & ~(15 << 27)) | (((local >> 4) & 15) << 27);
        t.z = (t.z & ~(15 << 27)) | (((local >> 8) & 15) << 27);
        q = pq[p];
        if (pmat) mm = pmat[p];
        if (tpa) {
            a = pa[p];
            b = pb[p];
            c = pc[p];
        }
    }
    tpt[g] = t;
    tpq[g] = q;
    if (tpm) tpm[g] = mm;
    if (tpa) {
        tpa[g] = a;
        tpb[g] = b;
        tpc[g] = c;
    }
}

// The padded (mu, lambda) slots of the current tile cut rewritten from pmat, laid out as
// k_pad_records lays them out (slot j of entry i: pmat[tileRec[i] + j] below the entry's
// record count, else zero).  Run after the per-tet materials change.
__global__ __launch_bounds__(256) void k_pad_pmat(int nentries, int rt, const int* __restrict__ tileRec,
                                                  const f2* __restrict__ pmat, f2* __restrict__ tpm) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (int64_t)nentries * rt) return;
    const int i = (int)(g / rt), j = (int)(g % rt);
    const int r0 = tileRec[i];
    const int cnt = i + 1 < nentries ? tileRec[i + 1] - r0 : 0;
    f2 mm = f2{0.f, 0.f};
    if (j < cnt) mm = pmat[r0 + j];
    tpm[g] = mm;
}

hipError_t launch_pad_pmat(const DevMesh& m, hipStream_t s) {
    if (!m.tpm || !m.pmat || m.tileEntries <= 0 || m.tileRecords <= 0) return hipSuccess;
    const int64_t total = (int64_t)m.tileEntries * m.tileRecords;
    hipLaunchKernelGGL(k_pad_pmat, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, m.tileEntries,
                       m.tileRecords, m.tileRec, m.pmat, m.tpm);
    return hipGetLastError();
}

void free_padded_records(DevMesh& m) {
    void* ps[] = {m.tpt, m.tpq, m.tpm, m.tpa, m.tpb, m.tpc};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    m.bytes -= m.padBytes;
    m.padBytes = 0;
    m.tpt = nullptr;
    m.tpq = nullptr;
    m.tpm = nullptr;
    m.tpa = nullptr;
    m.tpb = nullptr;
    m.tpc = nullptr;
}

// (Re)builds the padded record slots of the current tile cut; full adds the full records
// (pa, pb, pc) the corotated material reads.
hipError_t pad_records(DevMesh& m, bool full, hipStream_t s, std::string& err) {
    if (full) {
        hipError_t ef = ensure_full_records(m, s, err);
        if (ef != hipSuccess) return ef;
    }
    const int nentries = m.tileEntries;
    const size_t n = (size_t)std::max(nentries, 1) * (size_t)m.tileRecords;
    (void)hipStreamSynchronize(s);
    free_padded_records(m);
    hipError_t e = hipMalloc(&m.tpt, sizeof(i4) * n);
    if (e == hipSuccess) e = hipMalloc(&m.tpq, sizeof(f4) * n);
    if (e == hipSuccess && m.pmat) e = hipMalloc(&m.tpm, sizeof(f2) * n);
    if (e == hipSuccess && full) e = hipMalloc(&m.tpa, sizeof(f4) * n);
    if (e == hipSuccess && full) e = hipMalloc(&m.tpb, sizeof(f4) * n);
    if (e == hipSuccess && full) e = hipMalloc(&m.tpc, sizeof(float) * n);
    if (e == hipSuccess) {
        m.padBytes = (int64_t)n * (int64_t)(sizeof(i4) + sizeof(f4) + (m.tpm ? sizeof(f2) : 0) +
                                            (full ? 2 * sizeof(f4) + sizeof(float) : 0));
        m.bytes += m.padBytes;
    }
    if (e == hipSuccess && nentries > 0) {
        const int64_t total = (int64_t)nentries * m.tileRecords;
        hipLaunchKernelGGL(k_pad_records, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, nentries,
                           m.tileRecords, m.tileStart, m.tileRec, m.chunkBase, m.ptet, m.pq, m.pmat, m.pa, m.pb, m.pc,
                           m.tpt, m.tpq, m.tpm, m.tpa, m.tpb, m.tpc);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        err = std::string("padded tile records: ") + hipGetErrorString(e);
        free_padded_records(m);
    }
    return e;
}

hipError_t build_tiles(DevMesh& m, int records, int max_verts, hipStream_t s, std::string& err, bool full) {
    if (m.tileRecords == records && m.tileVerts == max_verts && m.tileStart) {
        if (full && !m.tpa) return pad_records(m, true, s, err);
        return hipSuccess;
    }
    if (!m.tiledOk) {
        err = "tiled kernel unavailable: a 4096-record chunk spans more than 4096 vertices";
        return hipErrorInvalidValue;
    }
    if (m.maxValence > records) {
        err = "records per tile (" + std::to_string(records) + ") below the maximum valence (" +
              std::to_string(m.maxValence) + ")";
        return hipErrorInvalidValue;
    }
    for (int c = 0; c < m.nColors; ++c)
        if ((int64_t)m.hostRow[m.colorSlotOff[c + 1]] - m.hostRow[m.colorSlotOff[c]] >= (int64_t)1 << 27) {
            err = "tiled kernel: a color holds 2^27 records or more (32-bit buffer offsets)";
            return hipErrorInvalidValue;
        }
    std::vector<int> starts;
    std::vector<int> base(m.nColors, 0), count(m.nColors, 0);
    for (int c = 0; c < m.nColors; ++c) {
        base[c] = (int)starts.size();
        const int e = m.colorSlotOff[c + 1];
        int v = m.colorSlotOff[c];
        while (v < e) {
            starts.push_back(v);
            const int r = m.hostRow[v];
            while (v < e && m.hostRow[v + 1] - r <= records && v - starts.back() < max_verts) ++v;
            if (v == starts.back()) {  // unreachable after the valence check above
                err = "tile cut failed";
                return hipErrorInvalidValue;
            }
            ++count[c];
        }
        starts.push_back(e);
    }
    std::vector<int> recs(starts.size());
    for (size_t i = 0; i < starts.size(); ++i) recs[i] = m.hostRow[starts[i]];
    if (m.tileStart) {
        (void)hipStreamSynchronize(s);
        (void)hipFree(m.tileStart);
        (void)hipFree(m.tileRec);
        m.tileStart = nullptr;
        m.tileRec = nullptr;
        m.tileRecords = 0;
        free_padded_records(m);
    }
    const size_t n = starts.empty() ? 1 : starts.size();
    hipError_t e = hipMalloc(&m.tileStart, sizeof(int) * n);
    if (e == hipSuccess) e = hipMalloc(&m.tileRec, sizeof(int) * n);
    if (e == hipSuccess && !starts.empty())
        e = hipMemcpyAsync(m.tileStart, starts.data(), sizeof(int) * starts.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && !starts.empty())
        e = hipMemcpyAsync(m.tileRec, recs.data(), sizeof(int) * recs.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        err = std::string("tile upload: ") + hipGetErrorString(e);
        if (m.tileStart) (void)hipFree(m.tileStart);
        if (m.tileRec) (void)hipFree(m.tileRec);
        m.tileStart = nullptr;
        m.tileRec = nullptr;
        return e;
    }
    m.colorTileBase = base;
    m.colorTileCount = count;
    if (m.dColorTileBase) (void)hipFree(m.dColorTileBase);
    m.dColorTileBase = nullptr;
    if (m.nColors > 0) {
        e = hipMalloc(&m.dColorTileBase, sizeof(int) * 2 * (size_t)m.nColors);
        if (e == hipSuccess) e = hipMemcpyAsync(m.dColorTileBase, base.data(), sizeof(int) * m.nColors, hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(m.dColorTileBase + m.nColors, count.data(), sizeof(int) * m.nColors, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            err = std::string("tile tables: ") + hipGetErrorString(e);
            return e;
        }
        m.dColorTileCount = m.dColorTileBase + m.nColors;
    }
    m.tileRecords = records;
    m.tileVerts = max_verts;
    m.tileEntries = (int)starts.size();
    e = pad_records(m, full, s, err);
    if (e != hipSuccess) {
        (void)hipFree(m.tileStart);
        (void)hipFree(m.tileRec);
        m.tileStart = nullptr;
        m.tileRec = nullptr;
        m.tileRecords = 0;
        m.tileVerts = 0;
    }
    return e;
}

void detach_tiles(DevMesh& m, TileCut& out) {
    out.tileStart = m.tileStart;
    out.tileRec = m.tileRec;
    out.tileRecords = m.tileRecords;
    out.tileVerts = m.tileVerts;
    out.tileEntries = m.tileEntries;
    out.tpt = m.tpt;
    out.tpq = m.tpq;
    out.tpm = m.tpm;
    out.tpa = m.tpa;
    out.tpb = m.tpb;
    out.tpc = m.tpc;
    out.padBytes = m.padBytes;
    out.colorTileBase.swap(m.colorTileBase);
    out.colorTileCount.swap(m.colorTileCount);
    out.dColorTileBase = m.dColorTileBase;
    out.dColorTileCount = m.dColorTileCount;
    m.dColorTileBase = m.dColorTileCount = nullptr;
    m.bytes -= m.padBytes;
    m.tileStart = m.tileRec = nullptr;
    m.tileRecords = m.tileVerts = m.tileEntries = 0;
    m.tpt = nullptr;
    m.tpq = nullptr;
    m.tpm = nullptr;
    m.tpa = m.tpb = nullptr;
    m.tpc = nullptr;
    m.padBytes = 0;
    m.colorTileBase.clear();
    m.colorTileCount.clear();
}

void free_tile_cut(TileCut& t) {
    void* ps[] = {t.tileStart, t.tileRec, t.tpt, t.tpq, t.tpm, t.tpa, t.tpb, t.tpc, t.dColorTileBase};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    t = TileCut();
}

void attach_tiles(DevMesh& m, TileCut& in, hipStream_t s) {
    (void)hipStreamSynchronize(s);
    TileCut cur;
    detach_tiles(m, cur);
    free_tile_cut(cur);
    m.tileStart = in.tileStart;
    m.tileRec = in.tileRec;
    m.tileRecords = in.tileRecords;
    m.tileVerts = in.tileVerts;
    m.tileEntries = in.tileEntries;
    m.tpt = in.tpt;
    m.tpq = in.tpq;
    m.tpm = in.tpm;
    m.tpa = in.tpa;
    m.tpb = in.tpb;
    m.tpc = in.tpc;
    m.padBytes = in.padBytes;
    m.bytes += in.padBytes;
    m.colorTileBase.swap(in.colorTileBase);
    m.colorTileCount.swap(in.colorTileCount);
    m.dColorTileBase = in.dColorTileBase;
    m.dColorTileCount = in.dColorTileCount;
    in = TileCut();
}

template <int LPV, int U, bool MAT, bool CONTACT, bool SAVE>
static hipError_t launch_color_t(const DevMesh& m, int color, int block, float sdt, float mu, float lambda,
                                 const ContactArgs& ct, f4* save, hipStream_t s) {
    const int slot0 = m.colorSlotOff[color];
    const int count = m.colorSlotOff[color + 1] - slot0;
    if (count <= 0) return hipSuccess;
    if (!m.pa) return hipErrorInvalidValue;  // full records not built (ensure_full_records)
    const int gpb = block / LPV;
    int nb = (count + gpb - 1) / gpb;
    nb = (nb + 7) & ~7;
    hipLaunchKernelGGL((k_color_solve<LPV, U, MAT, CONTACT, SAVE>), dim3(nb), dim3(block), 0, s, slot0, count, m.nV, m.pos,
                       m.inert, m.mass, m.row, m.ptet, m.pa, m.pb, m.pc, m.pmat, save, mu, lambda, sdt, ct);
    return hipGetLastError();
}

template <int LPV, int U>
static hipError_t launch_color_lu(const DevMesh& m, int color, int block, float sdt, float mu, float lambda,
                                  const ContactArgs& ct, f4* save, hipStream_t s) {
    const bool mat = m.perTetMat, con = ct.on != 0, sv = save != nullptr;
#define VBD_DISPATCH(MT, CT, SV) \
    if (mat == MT && con == CT && sv == SV) return launch_color_t<LPV, U, MT, CT, SV>(m, color, block, sdt, mu, lambda, ct, save, s);
    VBD_DISPATCH(false, false, false) VBD_DISPATCH(false, false, true)
    VBD_DISPATCH(false, true, false) VBD_DISPATCH(false, true, true)
    VBD_DISPATCH(true, false, false) VBD_DISPATCH(true, false, true)
    VBD_DISPATCH(true, true, false) VBD_DISPATCH(true, true, true)
#undef VBD_DISPATCH
    return hipErrorInvalidValue;
}

bool launch_config_supported(int lpv, int unroll, int block) {
    if (block < 64 || block > 256 || (block % 64) != 0) return false;
    if (lpv == 4 || lpv == 8 || lpv == 16) return unroll >= 1 && unroll <= 3;
    if (lpv == 32 || lpv == 64) return unroll == 1;
    return false;
}

hipError_t launch_color(const DevMesh& m, int color, int lpv, int unroll, int block, float sdt, float mu, float lambda,
                        const ContactArgs& ct, f4* save, hipStream_t s) {
    if (!launch_config_supported(lpv, unroll, block)) return hipErrorInvalidValue;
#define VBD_LU(L, UU) \
    if (lpv == L && unroll == UU) return launch_color_lu<L, UU>(m, color, block, sdt, mu, lambda, ct, save, s);
    VBD_LU(4, 1) VBD_LU(4, 2) VBD_LU(4, 3)
    VBD_LU(8, 1) VBD_LU(8, 2) VBD_LU(8, 3)
    VBD_LU(16, 1) VBD_LU(16, 2) VBD_LU(16, 3)
    VBD_LU(32, 1) VBD_LU(64, 1)
#undef VBD_LU
    return hipErrorInvalidValue;
}

template <int LPV, bool MAT, bool CONTACT, bool SAVE>
__global__ __launch_bounds__(256) void k_color_corot(int slot0, int count, int nslots, f4* __restrict__ pos,
                                                     const f4* __restrict__ inert, const float* __restrict__ mass,
                                                     const int* __restrict__ row, const i4* __restrict__ ptet,
                                                     const f4* __restrict__ pa, const f4* __restrict__ pb,
                                                     const float* __restrict__ pc, const f2* __restrict__ pmat,
                                                     f4* __restrict__ save, float mu, float lambda, float sdt,
                                                     ContactArgs ct) {
    const int gpb = blockDim.x / LPV;
    const int g = xcd_block() * gpb + (int)threadIdx.x / LPV;
    const int lane = (int)threadIdx.x % LPV;
    const bool valid = g < count;
    const int s = slot0 + g;
    Acc A = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    f4 xs = {0.f, 0.f, 0.f, 0.f};
    if (valid) {
        const __amdgpu_buffer_rsrc_t prs = pos_rsrc(pos, nslots);
        xs = pos[s];
        const int p1 = row[s + 1];
        for (int p = row[s] + lane; p < p1; p += LPV) {
            const i4 t = __builtin_nontemporal_load(&ptet[p]);
            const f4 a = __builtin_nontemporal_load(&pa[p]);
            const f4 b = __builtin_nontemporal_load(&pb[p]);
            const float c = __builtin_nontemporal_load(&pc[p]);
            float tmu = mu, tlam = lambda;
            if constexpr (MAT) {
                const f2 mm = __builtin_nontemporal_load(&pmat[p]);
                tmu = mm.x;
                tlam = mm.y;
            }
            const f4 ya = gather_pos(prs, t.x & SLOT_MASK), yb = gather_pos(prs, t.y & SLOT_MASK);
            const f4 yc = gather_pos(prs, t.z & SLOT_MASK);
            corot_accumulate(t.x < 0, a, b, c, __int_as_float(t.w), tmu, tlam, xs, ya, yb, yc, A);
        }
    }
    float h00 = gsum<LPV>(A.h00), h01 = gsum<LPV>(A.h01), h02 = gsum<LPV>(A.h02);
    float h11 = gsum<LPV>(A.h11), h12 = gsum<LPV>(A.h12), h22 = gsum<LPV>(A.h22);
    float r0 = gsum<LPV>(A.r0), r1 = gsum<LPV>(A.r1), r2 = gsum<LPV>(A.r2);
    if (!valid || lane != LPV - 1) return;
    const f4 xn = corot_vertex_solve<CONTACT>(xs, inert[s], mass[s], h00, h01, h02, h11, h12, h22, r0, r1, r2, sdt, ct);
    if constexpr (SAVE) save[s] = xs;
    pos[s] = xn;
}

template <int LPV, bool MAT, bool CONTACT, bool SAVE>
static hipError_t launch_corot_t(const DevMesh& m, int color, int block, float sdt, float mu, float lambda,
                                 const ContactArgs& ct, f4* save, hipStream_t s) {
    const int slot0 = m.colorSlotOff[color];
    const int count = m.colorSlotOff[color + 1] - slot0;
    if (count <= 0) return hipSuccess;
    if (!m.pa) return hipErrorInvalidValue;  // full records not built (ensure_full_records)
    const int gpb = block / LPV;
    int nb = (count + gpb - 1) / gpb;
    nb = (nb + 7) & ~7;
    hipLaunchKernelGGL((k_color_corot<LPV, MAT, CONTACT, SAVE>), dim3(nb), dim3(block), 0, s, slot0, count, m.nV,
                       m.pos, m.inert, m.mass, m.row, m.ptet, m.pa, m.pb, m.pc, m.pmat, save, mu, lambda, sdt, ct);
    return hipGetLastError();
}

template <int LPV>
static hipError_t launch_corot_l(const DevMesh& m, int color, int block, float sdt, float mu, float lambda,
                                 const ContactArgs& ct, f4* save, hipStream_t s) {
    const bool mat = m.perTetMat, con = ct.on != 0, sv = save != nullptr;
#define VBD_DISPATCH(MT, CT, SV) \
    if (mat == MT && con == CT && sv == SV) return launch_corot_t<LPV, MT, CT, SV>(m, color, block, sdt, mu, lambda, ct, save, s);
    VBD_DISPATCH(false, false, false) VBD_DISPATCH(false, false, true)
    VBD_DISPATCH(false, true, false) VBD_DISPATCH(false, true, true)
    VBD_DISPATCH(true, false, false) VBD_DISPATCH(true, false, true)
    VBD_DISPATCH(true, true, false) VBD_DISPATCH(true, true, true)
#undef VBD_DISPATCH
    return hipErrorInvalidValue;
}

hipError_t launch_color_corotated(const DevMesh& m, int color, int lpv, int block, float sdt, float mu, float lambda,
                                  const ContactArgs& ct, f4* save, hipStream_t s) {
    if (block < 64 || block > 256 || (block % 64) != 0) return hipErrorInvalidValue;
    switch (lpv) {
        case 8: return launch_corot_l<8>(m, color, block, sdt, mu, lambda, ct, save, s);
        case 16: return launch_corot_l<16>(m, color, block, sdt, mu, lambda, ct, save, s);
        case 32: return launch_corot_l<32>(m, color, block, sdt, mu, lambda, ct, save, s);
        default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------------------
// Differentiation kernels.
//
// With A = M/h^2 + K(x) (K = Hessian of E, plus the penalty-contact term k e_a e_a^T where a
// vertex penetrates), the adjoint of one implicit step solves A w = rhs.  The per-(vertex, tet)
// Hessian-vector product needs the directional derivative of P along dF = sum_n w_n q_n^T:
//   dP = mu (dF - dC) + lambda (C : dF) C + lambda (J - 1) dC,   dC = d cof(F)[dF]
// (Energy.cpp:76-88 and :123-127), and (K_t w)_k = V dP q_k.
// ---------------------------------------------------------------------------------------
enum PassMode { MODE_ADJ_SOLVE = 0, MODE_ADJ_RESID = 1, MODE_HVP = 2, MODE_FWD_RESID = 3 };

__device__ __forceinline__ void pair_hvp(const Kin& k, const bool corner0, const f4 a, const f4 b, const float c,
                                         const float V, const float tmu, const float tlam, const f4 ws, const f4 wa,
                                         const f4 wb, const f4 wc, float& o0, float& o1, float& o2) {
    const f4 base = corner0 ? ws : wa;
    const f4 first = corner0 ? wa : ws;
    const float d0x = first.x - base.x, d0y = first.y - base.y, d0z = first.z - base.z;
    const float d1x = wb.x - base.x, d1y = wb.y - base.y, d1z = wb.z - base.z;
    const float d2x = wc.x - base.x, d2y = wc.y - base.y, d2z = wc.z - base.z;
    const float G00 = d0x * a.x + d1x * a.w + d2x * b.z;
    const float G01 = d0x * a.y + d1x * b.x + d2x * b.w;
    const float G02 = d0x * a.z + d1x * b.y + d2x * c;
    const float G10 = d0y * a.x + d1y * a.w + d2y * b.z;
    const float G11 = d0y * a.y + d1y * b.x + d2y * b.w;
    const float G12 = d0y * a.z + d1y * b.y + d2y * c;
    const float G20 = d0z * a.x + d1z * a.w + d2z * b.z;
    const float G21 = d0z * a.y + d1z * b.x + d2z * b.w;
    const float G22 = d0z * a.z + d1z * b.y + d2z * c;
    const float dJ = k.C00 * G00 + k.C01 * G01 + k.C02 * G02 + k.C10 * G10 + k.C11 * G11 + k.C12 * G12 +
                     k.C20 * G20 + k.C21 * G21 + k.C22 * G22;
    // cofactor derivative (Energy.cpp:76-88) with dF = G
    const float dC00 = G11 * k.F22 + k.F11 * G22 - G12 * k.F21 - k.F12 * G21;
    const float dC01 = G12 * k.F20 + k.F12 * G20 - G10 * k.F22 - k.F10 * G22;
    const float dC02 = G10 * k.F21 + k.F10 * G21 - G11 * k.F20 - k.F11 * G20;
    const float dC10 = G21 * k.F02 + k.F21 * G02 - G22 * k.F01 - k.F22 * G01;
    const float dC11 = G22 * k.F00 + k.F22 * G00 - G20 * k.F02 - k.F20 * G02;
    const float dC12 = G20 * k.F01 + k.F20 * G01 - G21 * k.F00 - k.F21 * G00;
    const float dC20 = G01 * k.F12 + k.F01 * G12 - G02 * k.F11 - k.F02 * G11;
    const float dC21 = G02 * k.F10 + k.F02 * G10 - G00 * k.F12 - k.F00 * G12;
    const float dC22 = G00 * k.F11 + k.F00 * G11 - G01 * k.F10 - k.F01 * G10;
    const float ld = tlam * dJ;
    const float aj = tlam * (k.J - 1.0f);
    const float q0 = k.q0, q1 = k.q1, q2 = k.q2;
#define VBD_DP(r, c) (tmu * (G##r##c - dC##r##c) + ld * k.C##r##c + aj * dC##r##c)
    o0 = V * (VBD_DP(0, 0) * q0 + VBD_DP(0, 1) * q1 + VBD_DP(0, 2) * q2);
    o1 = V * (VBD_DP(1, 0) * q0 + VBD_DP(1, 1) * q1 + VBD_DP(1, 2) * q2);
    o2 = V * (VBD_DP(2, 0) * q0 + VBD_DP(2, 1) * q1 + VBD_DP(2, 2) * q2);
#undef VBD_DP
}

// One pass over free vertices [slot0, slot0 + count) with LPV lanes per vertex.
//   MODE_ADJ_SOLVE:  block Gauss-Seidel update of w for one color: r = rhs - (A w)_s,
//                    w_s += D_ss^-1 r with D_ss the exact 3x3 diagonal block of A; SAVE keeps the
//                    pre-update w_s for Chebyshev acceleration.
//   MODE_ADJ_RESID:  out[s] = rhs_s - (A w)_s,  r2[s] = |out[s]|^2
//   MODE_HVP:        out[s] = (A w)_s
//   MODE_FWD_RESID:  out[s] = grad G = m/h^2 (x_s - y_s) + sum_t g + contact,  r2[s] = |.|^2
template <int MODE, int LPV, bool MAT, bool CONTACT, bool SAVE>
__global__ __launch_bounds__(256) void k_vertex_pass(int slot0, int count, int nslots, const f4* __restrict__ pos,
                                                     const f4* __restrict__ inert, const float* __restrict__ mass,
                                                     const int* __restrict__ row, const i4* __restrict__ ptet,
                                                     const f4* __restrict__ pa, const f4* __restrict__ pb,
                                                     const float* __restrict__ pc, const f2* __restrict__ pmat,
                                                     f4* __restrict__ w, const f4* __restrict__ rhs,
                                                     f4* __restrict__ out, double* __restrict__ r2,
                                                     f4* __restrict__ save, float mu, float lambda, float sdt,
                                                     ContactArgs ct) {
    const int gpb = blockDim.x / LPV;
    const int g = xcd_block() * gpb + (int)threadIdx.x / LPV;
    const int lane = (int)threadIdx.x % LPV;
    const bool valid = g < count;
    const int s = slot0 + g;
    constexpr bool NEED_W = MODE != MODE_FWD_RESID;

    Acc A = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float kw0 = 0.f, kw1 = 0.f, kw2 = 0.f;
    f4 xs = {0.f, 0.f, 0.f, 0.f}, ws = {0.f, 0.f, 0.f, 0.f};
    if (valid) {
        const __amdgpu_buffer_rsrc_t prs = pos_rsrc(pos, nslots);
        __amdgpu_buffer_rsrc_t wrs = prs;
        if constexpr (NEED_W) {
            wrs = pos_rsrc(w, nslots);
            ws = w[s];
        }
        xs = pos[s];
        const int p0 = row[s], p1 = row[s + 1];
        for (int p = p0 + lane; p < p1; p += LPV) {
            const i4 t = ptet[p];
            const f4 a = pa[p];
            const f4 b = pb[p];
            const float c = pc[p];
            float tmu = mu, tlam = lambda;
            if constexpr (MAT) {
                const f2 mm = pmat[p];
                tmu = mm.x;
                tlam = mm.y;
            }
            const int sa = t.x & SLOT_MASK, sb = t.y & SLOT_MASK, sc = t.z & SLOT_MASK;
            const bool c0 = t.x < 0;
            const f4 ya = gather_pos(prs, sa), yb = gather_pos(prs, sb), yc = gather_pos(prs, sc);
            const float V = __int_as_float(t.w);
            if constexpr (MODE == MODE_FWD_RESID) {
                pair_accumulate(c0, a, b, c, V, tmu, tlam, xs, ya, yb, yc, A);
            } else {
                const f4 wa = gather_pos(wrs, sa), wb = gather_pos(wrs, sb), wc = gather_pos(wrs, sc);
                const Kin k = pair_kinematics(c0, a, b, c, xs, ya, yb, yc);
                float o0, o1, o2;
                pair_hvp(k, c0, a, b, c, V, tmu, tlam, ws, wa, wb, wc, o0, o1, o2);
                kw0 += o0;
                kw1 += o1;
                kw2 += o2;
                if constexpr (MODE == MODE_ADJ_SOLVE) {
                    const float Cq0 = k.C00 * k.q0 + k.C01 * k.q1 + k.C02 * k.q2;
                    const float Cq1 = k.C10 * k.q0 + k.C11 * k.q1 + k.C12 * k.q2;
                    const float Cq2 = k.C20 * k.q0 + k.C21 * k.q1 + k.C22 * k.q2;
                    const float vm = V * tmu * (k.q0 * k.q0 + k.q1 * k.q1 + k.q2 * k.q2);
                    const float vl = V * tlam;
                    A.h00 += vm + vl * Cq0 * Cq0;
                    A.h01 += vl * Cq0 * Cq1;
                    A.h02 += vl * Cq0 * Cq2;
                    A.h11 += vm + vl * Cq1 * Cq1;
                    A.h12 += vl * Cq1 * Cq2;
                    A.h22 += vm + vl * Cq2 * Cq2;
                }
            }
        }
    }
    kw0 = gsum<LPV>(kw0); kw1 = gsum<LPV>(kw1); kw2 = gsum<LPV>(kw2);
    float h00 = 0.f, h01 = 0.f, h02 = 0.f, h11 = 0.f, h12 = 0.f, h22 = 0.f, r0 = 0.f, r1 = 0.f, r2v = 0.f;
    if constexpr (MODE == MODE_ADJ_SOLVE) {
        h00 = gsum<LPV>(A.h00); h01 = gsum<LPV>(A.h01); h02 = gsum<LPV>(A.h02);
        h11 = gsum<LPV>(A.h11); h12 = gsum<LPV>(A.h12); h22 = gsum<LPV>(A.h22);
    }
    if constexpr (MODE == MODE_FWD_RESID) {
        r0 = gsum<LPV>(A.r0); r1 = gsum<LPV>(A.r1); r2v = gsum<LPV>(A.r2);
    }
    if (!valid || lane != LPV - 1) return;
    const float m = mass[s];
    const float ms = m * sdt;
    // penalty contact: Hessian k e_a e_a^T and gradient -k (plane - x_a) e_a while penetrating
    float ck = 0.f, cd = 0.f;
    int cax = 0;
    if constexpr (CONTACT) {
        cax = ct.axis;
        const float xa = cax == 0 ? xs.x : (cax == 1 ? xs.y : xs.z);
        cd = ct.plane - xa;
        if (cd > 0.f) ck = ct.k;
        else cd = 0.f;
    }
    if constexpr (MODE == MODE_FWD_RESID) {
        const f4 ys = inert[s];
        f4 o;
        o.x = (m * (xs.x - ys.x)) * sdt - r0;
        o.y = (m * (xs.y - ys.y)) * sdt - r1;
        o.z = (m * (xs.z - ys.z)) * sdt - r2v;
        o.w = 0.f;
        if constexpr (CONTACT) {
            if (cax == 0) o.x -= ck * cd;
            else if (cax == 1) o.y -= ck * cd;
            else o.z -= ck * cd;
        }
        out[s] = o;
        r2[s] = (double)o.x * o.x + (double)o.y * o.y + (double)o.z * o.z;
        return;
    }
    // (A w)_s
    float aw0 = ms * ws.x + kw0, aw1 = ms * ws.y + kw1, aw2 = ms * ws.z + kw2;
    if constexpr (CONTACT) {
        if (cax == 0) aw0 += ck * ws.x;
        else if (cax == 1) aw1 += ck * ws.y;
        else aw2 += ck * ws.z;
    }
    if constexpr (MODE == MODE_HVP) {
        f4 o;
        o.x = aw0; o.y = aw1; o.z = aw2; o.w = 0.f;
        out[s] = o;
        return;
    }
    const f4 b4 = rhs[s];
    const float e0 = b4.x - aw0, e1 = b4.y - aw1, e2 = b4.z - aw2;
    if constexpr (MODE == MODE_ADJ_RESID) {
        f4 o;
        o.x = e0; o.y = e1; o.z = e2; o.w = 0.f;
        out[s] = o;
        r2[s] = (double)e0 * e0 + (double)e1 * e1 + (double)e2 * e2;
        return;
    }
    // MODE_ADJ_SOLVE
    h00 += ms; h11 += ms; h22 += ms;
    if constexpr (CONTACT) {
        if (cax == 0) h00 += ck;
        else if (cax == 1) h11 += ck;
        else h22 += ck;
    }
    const float det = h00 * (h11 * h22 - h12 * h12) - h01 * (h01 * h22 - h12 * h02) + h02 * (h01 * h12 - h11 * h02);
    f4 wn = ws;
    if (fabsf(det) > 1e-8f) {
        const float inv = 1.0f / det;
        const float a00 = h11 * h22 - h12 * h12, a01 = h02 * h12 - h01 * h22, a02 = h01 * h12 - h02 * h11;
        const float a11 = h00 * h22 - h02 * h02, a12 = h02 * h01 - h00 * h12, a22 = h00 * h11 - h01 * h01;
        wn.x = ws.x + inv * (a00 * e0 + a01 * e1 + a02 * e2);
        wn.y = ws.y + inv * (a01 * e0 + a11 * e1 + a12 * e2);
        wn.z = ws.z + inv * (a02 * e0 + a12 * e1 + a22 * e2);
    }
    if constexpr (SAVE) save[s] = ws;
    w[s] = wn;
}

template <int MODE, int LPV, bool MAT, bool CONTACT, bool SAVE>
static hipError_t launch_pass_t(const DevMesh& m, int slot0, int count, int block, f4* w, const f4* rhs, f4* out,
                                double* r2, f4* save, float mu, float lambda, float sdt, const ContactArgs& ct,
                                hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (!m.pa) return hipErrorInvalidValue;  // full records not built (ensure_full_records)
    const int gpb = block / LPV;
    int nb = (count + gpb - 1) / gpb;
    nb = (nb + 7) & ~7;
    hipLaunchKernelGGL((k_vertex_pass<MODE, LPV, MAT, CONTACT, SAVE>), dim3(nb), dim3(block), 0, s, slot0, count, m.nV,
                       m.pos, m.inert, m.mass, m.row, m.ptet, m.pa, m.pb, m.pc, m.pmat, w, rhs, out, r2, save, mu,
                       lambda, sdt, ct);
    return hipGetLastError();
}

template <int MODE, int LPV>
static hipError_t launch_pass_l(const DevMesh& m, int slot0, int count, int block, f4* w, const f4* rhs, f4* out,
                                double* r2, f4* save, float mu, float lambda, float sdt, const ContactArgs& ct,
                                hipStream_t s) {
    const bool mat = m.perTetMat, con = ct.on != 0, sv = save != nullptr;
#define VBD_PD(MT, CT, SV) \
    if (mat == MT && con == CT && sv == SV) \
        return launch_pass_t<MODE, LPV, MT, CT, SV>(m, slot0, count, block, w, rhs, out, r2, save, mu, lambda, sdt, ct, s);
    if constexpr (MODE == MODE_ADJ_SOLVE) {
        VBD_PD(false, false, true) VBD_PD(false, true, true) VBD_PD(true, false, true) VBD_PD(true, true, true)
    }
    VBD_PD(false, false, false) VBD_PD(false, true, false) VBD_PD(true, false, false) VBD_PD(true, true, false)
#undef VBD_PD
    return hipErrorInvalidValue;
}

hipError_t launch_pass(int mode, const DevMesh& m, int slot0, int count, int lpv, int block, f4* w, const f4* rhs,
                       f4* out, double* r2, f4* save, float mu, float lambda, float sdt, const ContactArgs& ct,
                       hipStream_t s) {
    if (block < 64 || block > 256 || (block % 64) != 0 || (block % lpv) != 0) return hipErrorInvalidValue;
#define VBD_PM(MD, L) \
    if (mode == MD && lpv == L) return launch_pass_l<MD, L>(m, slot0, count, block, w, rhs, out, r2, save, mu, lambda, sdt, ct, s);
    VBD_PM(MODE_ADJ_SOLVE, 8) VBD_PM(MODE_ADJ_SOLVE, 16) VBD_PM(MODE_ADJ_SOLVE, 32)
    VBD_PM(MODE_ADJ_RESID, 8) VBD_PM(MODE_ADJ_RESID, 16) VBD_PM(MODE_ADJ_RESID, 32)
    VBD_PM(MODE_HVP, 8) VBD_PM(MODE_HVP, 16) VBD_PM(MODE_HVP, 32)
    VBD_PM(MODE_FWD_RESID, 8) VBD_PM(MODE_FWD_RESID, 16) VBD_PM(MODE_FWD_RESID, 32)
#undef VBD_PM
    return hipErrorInvalidValue;
}

// Deterministic sum of n doubles: one block, fixed strided partition, fixed tree.
__global__ __launch_bounds__(1024) void k_sum_fixed(int n, const double* __restrict__ v, double* __restrict__ out) {
    __shared__ double sh[1024];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) acc += v[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

__global__ __launch_bounds__(256) void k_norm2_f4(int n, const f4* __restrict__ v, double* __restrict__ r2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f4 a = v[i];
    r2[i] = (double)a.x * a.x + (double)a.y * a.y + (double)a.z * a.z;
}

hipError_t launch_norm2_f4(int n, const f4* v, double* r2, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_norm2_f4, dim3((n + 255) / 256), dim3(256), 0, s, n, v, r2);
    return hipGetLastError();
}

// Stage one of the two-stage fixed-order sum: block b sums chunk b of ceil(n / SUM_STAGE).
__global__ __launch_bounds__(256) void k_sum_stage(int n, const double* __restrict__ v, double* __restrict__ stage) {
    __shared__ double sh[256];
    const int chunk = (n + SUM_STAGE - 1) / SUM_STAGE;
    const int b0 = (int)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
    double acc = 0.0;
    for (int i = b0 + (int)threadIdx.x; i < b1; i += 256) acc += v[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) stage[blockIdx.x] = sh[0];
}

hipError_t launch_sum_fixed(int n, const double* v, double* out, hipStream_t s, double* stage) {
    if (stage && n > SUM_TWO_STAGE) {
        hipLaunchKernelGGL(k_sum_stage, dim3(SUM_STAGE), dim3(256), 0, s, n, v, stage);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(1024), 0, s, SUM_STAGE, stage, out);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(1024), 0, s, n, v, out);
    return hipGetLastError();
}

// |m_s g|^2 per free slot (scale of the forward convergence test)
__global__ __launch_bounds__(256) void k_gravity_load2(int n, const float* __restrict__ mass, float gx, float gy,
                                                       float gz, double* __restrict__ r2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = mass[i];
    r2[i] = m * m * ((double)gx * gx + (double)gy * gy + (double)gz * gz);
}

hipError_t launch_gravity_load2(const DevMesh& m, const float g[3], double* r2, hipStream_t s) {
    if (m.nFree == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gravity_load2, dim3((m.nFree + 255) / 256), dim3(256), 0, s, m.nFree, m.mass, g[0], g[1],
                       g[2], r2);
    return hipGetLastError();
}

// Per-tet parameter gradients -(d^2E / dx dtheta)^T w in caller tet order:
//   dmu_t = -V (F - C) : dF(w),   dlambda_t = -V (J - 1) C : dF(w)
__global__ __launch_bounds__(256) void k_param_grad(int nT, const i4* __restrict__ tets, const f4* __restrict__ pos, const f4* __restrict__ w,
                                                    const float* __restrict__ dminv, const float* __restrict__ vol,
                                                    float* __restrict__ dmu, float* __restrict__ dlam) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const i4 tt = tets[t];
    const int sl[4] = {tt.x, tt.y, tt.z, tt.w};  // device slots
    const f4 x0 = pos[sl[0]], x1 = pos[sl[1]], x2 = pos[sl[2]], x3 = pos[sl[3]];
    const f4 w0 = w[sl[0]], w1 = w[sl[1]], w2 = w[sl[2]], w3 = w[sl[3]];
    const float* D = dminv + 9 * (size_t)t;
    const float ds[3][3] = {{x1.x - x0.x, x2.x - x0.x, x3.x - x0.x},
                            {x1.y - x0.y, x2.y - x0.y, x3.y - x0.y},
                            {x1.z - x0.z, x2.z - x0.z, x3.z - x0.z}};
    const float dw[3][3] = {{w1.x - w0.x, w2.x - w0.x, w3.x - w0.x},
                            {w1.y - w0.y, w2.y - w0.y, w3.y - w0.y},
                            {w1.z - w0.z, w2.z - w0.z, w3.z - w0.z}};
    float F[3][3], G[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            F[r][c] = ds[r][0] * D[c] + ds[r][1] * D[3 + c] + ds[r][2] * D[6 + c];
            G[r][c] = dw[r][0] * D[c] + dw[r][1] * D[3 + c] + dw[r][2] * D[6 + c];
        }
    float C[3][3];
    C[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
    C[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
    C[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
    C[1][0] = F[2][1] * F[0][2] - F[2][2] * F[0][1];
    C[1][1] = F[2][2] * F[0][0] - F[2][0] * F[0][2];
    C[1][2] = F[2][0] * F[0][1] - F[2][1] * F[0][0];
    C[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
    C[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
    C[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    const float J = F[0][0] * C[0][0] + F[0][1] * C[0][1] + F[0][2] * C[0][2];
    float fc = 0.f, cg = 0.f;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            fc += (F[r][c] - C[r][c]) * G[r][c];
            cg += C[r][c] * G[r][c];
        }
    dmu[t] = -vol[t] * fc;
    dlam[t] = -vol[t] * (J - 1.0f) * cg;
}

hipError_t launch_param_grad(const DevMesh& m, const f4* w, float* dmu, float* dlam, hipStream_t s) {
    if (m.nT == 0) return hipSuccess;
    hipLaunchKernelGGL(k_param_grad, dim3((m.nT + 255) / 256), dim3(256), 0, s, m.nT, m.tets, m.pos, w, m.dminv,
                       m.vol, dmu, dlam);
    return hipGetLastError();
}

// Rewrite the per-record (mu, lambda) stream from per-tet arrays (same record order as the build).
__global__ __launch_bounds__(256) void k_update_pmat(int nFree, const int* __restrict__ perm,
                                                     const int* __restrict__ off, const int* __restrict__ entry,
                                                     const int* __restrict__ row, const float* __restrict__ tmu,
                                                     const float* __restrict__ tlam, f2* __restrict__ pmat) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nFree) return;
    const int v = perm[s];
    int p = row[s];
    for (int e = off[v]; e < off[v + 1]; ++e, ++p) {
        const int t = entry[e] >> 2;
        f2 mm;
        mm.x = tmu[t];
        mm.y = tlam[t];
        pmat[p] = mm;
    }
}

hipError_t launch_update_pmat(const DevMesh& m, hipStream_t s) {
    if (m.nFree == 0 || !m.perTetMat) return hipSuccess;
    hipLaunchKernelGGL(k_update_pmat, dim3((m.nFree + 255) / 256), dim3(256), 0, s, m.nFree, m.perm, m.adjOff,
                       m.adjEntry, m.row, m.tmu, m.tlam, m.pmat);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_blend_range(int n, f4* __restrict__ v, const f4* __restrict__ pp,
                                                     float omega) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f4 x = v[i];
    const f4 q = pp[i];
    f4 r = x;
    r.x = omega * (x.x - q.x) + q.x;
    r.y = omega * (x.y - q.y) + q.y;
    r.z = omega * (x.z - q.z) + q.z;
    v[i] = r;
}

hipError_t launch_blend_vec(int n, f4* v, const f4* pp, float omega, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_blend_range, dim3((n + 255) / 256), dim3(256), 0, s, n, v, pp, omega);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Chebyshev blend over the free slots: x = omega (x - x_pp) + x_pp   (VBD.cpp:290-300)
// ---------------------------------------------------------------------------------------
// The Chebyshev step (VBD.cpp:290-300): omega (x - x_pp) + x_pp, rounded after the product and
// after the sum as the reference's CPU code is (no contraction).
// x = omega (x - x_pp) + x_pp over the free slots (VBD.cpp:290-300).  With keep, the blended x
// also replaces x_pp: it is the next sweep's starting position, which that sweep's blend needs two
// iterations later, so the next sweep's color launches need not save it (their per-vertex stores
// come from one lane in four; here they are one coalesced store per vertex).
__global__ __launch_bounds__(256) void k_blend(int n, f4* __restrict__ pos, f4* __restrict__ pp, float omega,
                                               int keep) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f4 x = pos[i];
    const f4 q = pp[i];
    f4 r = x;
    r.x = omega * (x.x - q.x) + q.x;
    r.y = omega * (x.y - q.y) + q.y;
    r.z = omega * (x.z - q.z) + q.z;
    pos[i] = r;
    if (keep) pp[i] = r;
}

hipError_t launch_blend(const DevMesh& m, float omega, f4* pp, hipStream_t s, bool keep) {
    if (m.nFree == 0) return hipSuccess;
    hipLaunchKernelGGL(k_blend, dim3((m.nFree + 255) / 256), dim3(256), 0, s, m.nFree, m.pos, pp, omega, keep ? 1 : 0);
    return hipGetLastError();
}

// forwardStep (VBD.cpp:103-112): x_prev = x; free: v += h g, y = x + h v; pinned: y = x.
__global__ __launch_bounds__(256) void k_predict(int n, int nFree, const f4* __restrict__ pos, f4* __restrict__ vel,
                                                 f4* __restrict__ xprev, f4* __restrict__ inert, float h, float gx,
                                                 float gy, float gz) {
#pragma clang fp contract(off)  // product and sum rounded separately, as the reference's CPU code
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f4 x = pos[i];
    xprev[i] = x;
    f4 y = x;
    if (i < nFree) {
        f4 v = vel[i];
        v.x += h * gx;
        v.y += h * gy;
        v.z += h * gz;
        vel[i] = v;
        y.x = x.x + h * v.x;
        y.y = x.y + h * v.y;
        y.z = x.z + h * v.z;
    }
    inert[i] = y;
}

hipError_t launch_predict(const DevMesh& m, float h, const float g[3], hipStream_t s) {
    if (m.nV == 0) return hipSuccess;
    hipLaunchKernelGGL(k_predict, dim3((m.nV + 255) / 256), dim3(256), 0, s, m.nV, m.nFree, m.pos, m.vel, m.xprev,
                       m.inert, h, g[0], g[1], g[2]);
    return hipGetLastError();
}

// updateVelocity (TetMesh.h:305-311): free: v = (x - x_prev) / h * damping.
__global__ __launch_bounds__(256) void k_velocity(int nFree, const f4* __restrict__ pos,
                                                  const f4* __restrict__ xprev, f4* __restrict__ vel, float h,
                                                  float damp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nFree) return;
    const f4 x = pos[i], xp = xprev[i];
    f4 v = vel[i];
    v.x = ((x.x - xp.x) / h) * damp;
    v.y = ((x.y - xp.y) / h) * damp;
    v.z = ((x.z - xp.z) / h) * damp;
    vel[i] = v;
}

hipError_t launch_velocity(const DevMesh& m, float h, float damping, hipStream_t s) {
    if (m.nFree == 0) return hipSuccess;
    hipLaunchKernelGGL(k_velocity, dim3((m.nFree + 255) / 256), dim3(256), 0, s, m.nFree, m.pos, m.xprev, m.vel, h,
                       damping);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_copy_f4(int n, const f4* __restrict__ src, f4* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

hipError_t launch_copy_inertia(const DevMesh& m, hipStream_t s) {
    if (m.nFree == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_f4, dim3((m.nFree + 255) / 256), dim3(256), 0, s, m.nFree, m.inert, m.pos);
    return hipGetLastError();
}

hipError_t launch_copy_f4(int n, const f4* src, f4* dst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_f4, dim3((n + 255) / 256), dim3(256), 0, s, n, src, dst);
    return hipGetLastError();
}

// caller layout (3 floats per vertex, caller numbering) <-> slot layout (f4 per slot)
__global__ __launch_bounds__(256) void k_scatter_pos(int n, const float* __restrict__ src3, const int* __restrict__ inv,
                                                     f4* __restrict__ dst) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    f4 r;
    r.x = src3[3 * (size_t)v];
    r.y = src3[3 * (size_t)v + 1];
    r.z = src3[3 * (size_t)v + 2];
    r.w = 0.f;
    dst[inv[v]] = r;
}

__global__ __launch_bounds__(256) void k_gather_pos(int n, const f4* __restrict__ src, const int* __restrict__ inv,
                                                    float* __restrict__ dst3) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const f4 r = src[inv[v]];
    dst3[3 * (size_t)v] = r.x;
    dst3[3 * (size_t)v + 1] = r.y;
    dst3[3 * (size_t)v + 2] = r.z;
}

hipError_t launch_scatter_pos(int n, const float* src3, const int* inv, f4* dst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_pos, dim3((n + 255) / 256), dim3(256), 0, s, n, src3, inv, dst);
    return hipGetLastError();
}

hipError_t launch_gather_pos(int n, const f4* src, const int* inv, float* dst3, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_pos, dim3((n + 255) / 256), dim3(256), 0, s, n, src, inv, dst3);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Energies, summed in double with a fixed order (per-block partial sums, then k_sum_fixed):
//   ENERGY_LOG_NH:    the reference's diagnostic (VBD.cpp:258-288 with NeoHookean.cpp:6-20),
//                     V0 [mu/2 (|F|^2 - 3) - mu ln J + lambda/2 (ln J)^2], J = |det F| (the
//                     product of the non-negative singular values);
//   ENERGY_STABLE_NH: the elastic part of the incremental potential the solver minimises,
//                     V0 [mu/2 (|F|^2 - 3) - mu (J - 1) + lambda/2 (J - 1)^2] (Energy.cpp:93-97);
//   ENERGY_COROTATED: V0 [mu |F - R|^2 + lambda/2 tr(R^T F - I)^2], R from the corotated
//                     kernel's 5 polar steps, whose gradient is that kernel's P;
// and the inertia part sum over free vertices of m/(2h^2) |x - y|^2 (+ penalty k/2 d^2).
// ---------------------------------------------------------------------------------------
constexpr int ENERGY_LOG_NH = 0, ENERGY_STABLE_NH = 1, ENERGY_COROTATED = 2;

__device__ __forceinline__ double block_sum_256(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    return red[0];
}

template <int MODE>
__global__ __launch_bounds__(256) void k_tet_energy(int nT, const i4* __restrict__ tets, const f4* __restrict__ pos,
                                                    const float* __restrict__ dminv,
                                                    const float* __restrict__ vol, const float* __restrict__ tmu,
                                                    const float* __restrict__ tlam, float mu, float lambda,
                                                    double* __restrict__ partial) {
    __shared__ double red[256];
    // The block's 256 DmInv (36 B each, not 16-B aligned per tet) are staged through LDS with
    // 16-byte coalesced loads (the block's segment starts at a multiple of 9216 bytes).  The
    // corner gathers are issued first and stay in flight across the LDS-only barrier.
    __shared__ f4 sD4[256 * 9 / 4];
    float* sD = reinterpret_cast<float*>(sD4);
    const int tid = (int)threadIdx.x;
    const int b0 = (int)blockIdx.x * 256;
    const int t = b0 + tid;
    const int nf = 9 * min(256, nT - b0);
    const float* Dg = dminv + 9 * (size_t)b0;
    f4 st[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (tid + 256 * k < nf / 4) st[k] = reinterpret_cast<const f4*>(Dg)[tid + 256 * k];
    const float tail = (nf / 4) * 4 + tid < nf ? Dg[(nf / 4) * 4 + tid] : 0.f;
    const i4 tt = tets[min(t, nT - 1)];
    const f4 x0 = pos[tt.x], x1 = pos[tt.y], x2 = pos[tt.z], x3 = pos[tt.w];  // device slots
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (tid + 256 * k < nf / 4) sD4[tid + 256 * k] = st[k];
    if ((nf / 4) * 4 + tid < nf) sD[(nf / 4) * 4 + tid] = tail;
    lds_barrier();
    double e = 0.0;
    if (t < nT) {
        const float* D = sD + 9 * tid;
        // F = Ds DmInv in double: near rest |F|^2 - 3 and J - 1 cancel to O(strain)
        const double ds[3][3] = {{(double)x1.x - x0.x, (double)x2.x - x0.x, (double)x3.x - x0.x},
                                 {(double)x1.y - x0.y, (double)x2.y - x0.y, (double)x3.y - x0.y},
                                 {(double)x1.z - x0.z, (double)x2.z - x0.z, (double)x3.z - x0.z}};
        double F[3][3];
        double tr = 0.0;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                F[r][c] = ds[r][0] * D[c] + ds[r][1] * D[3 + c] + ds[r][2] * D[6 + c];
                tr += F[r][c] * F[r][c];
            }
        const double det = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) -
                           F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
                           F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
        const double m_ = tmu ? tmu[t] : mu, l_ = tlam ? tlam[t] : lambda;
        if constexpr (MODE == ENERGY_LOG_NH) {
            const double lj = log(fabs(det));
            e = (0.5 * m_ * (tr - 3.0) - m_ * lj + 0.5 * l_ * lj * lj) * vol[t];
        } else if constexpr (MODE == ENERGY_COROTATED) {
            double R[3][3], C[3][3];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) R[r][c] = F[r][c];
            for (int it = 0; it < 5; ++it) {
                const double d = R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1]) -
                                 R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0]) +
                                 R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]);
                if (fabs(d) < 1e-8) break;
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c)
                        C[r][c] = R[(r + 1) % 3][(c + 1) % 3] * R[(r + 2) % 3][(c + 2) % 3] -
                                  R[(r + 1) % 3][(c + 2) % 3] * R[(r + 2) % 3][(c + 1) % 3];
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) R[r][c] = 0.5 * (R[r][c] + C[r][c] / d);
            }
            double fr = 0.0, rf = 0.0;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    fr += (F[r][c] - R[r][c]) * (F[r][c] - R[r][c]);
                    rf += R[r][c] * F[r][c];
                }
            e = (m_ * fr + 0.5 * l_ * (rf - 3.0) * (rf - 3.0)) * vol[t];
        } else {
            e = (0.5 * m_ * (tr - 3.0) - m_ * (det - 1.0) + 0.5 * l_ * (det - 1.0) * (det - 1.0)) * vol[t];
        }
    }
    const double sum = block_sum_256(e, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = sum;
}

__global__ __launch_bounds__(256) void k_inertia_energy(int nFree, const f4* __restrict__ pos,
                                                        const f4* __restrict__ inert, const float* __restrict__ mass,
                                                        double sdt, ContactArgs ct, double* __restrict__ partial) {
    __shared__ double red[256];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (s < nFree) {
        const f4 x = pos[s], y = inert[s];
        const double dx = (double)x.x - y.x, dy = (double)x.y - y.y, dz = (double)x.z - y.z;
        e = 0.5 * mass[s] * sdt * (dx * dx + dy * dy + dz * dz);
        if (ct.on) {
            const double xa = ct.axis == 0 ? x.x : (ct.axis == 1 ? x.y : x.z);
            const double d = (double)ct.plane - xa;
            if (d > 0.0) e += 0.5 * (double)ct.k * d * d;
        }
    }
    const double sum = block_sum_256(e, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = sum;
}

int energy_partials(const DevMesh& m) { return std::max(1, std::max((m.nT + 255) / 256, (m.nFree + 255) / 256)); }

hipError_t launch_energy(const DevMesh& m, int mode, float mu, float lambda, double* partial, double* out,
                         hipStream_t s) {
    const int nb = (m.nT + 255) / 256;
    if (nb == 0) return launch_sum_fixed(0, partial, out, s);
    if (mode == ENERGY_LOG_NH)
        hipLaunchKernelGGL(k_tet_energy<ENERGY_LOG_NH>, dim3(nb), dim3(256), 0, s, m.nT, m.tets, m.pos,
                           m.dminv, m.vol, m.tmu, m.tlam, mu, lambda, partial);
    else if (mode == ENERGY_COROTATED)
        hipLaunchKernelGGL(k_tet_energy<ENERGY_COROTATED>, dim3(nb), dim3(256), 0, s, m.nT, m.tets, m.pos,
                           m.dminv, m.vol, m.tmu, m.tlam, mu, lambda, partial);
    else
        hipLaunchKernelGGL(k_tet_energy<ENERGY_STABLE_NH>, dim3(nb), dim3(256), 0, s, m.nT, m.tets, m.pos,
                           m.dminv, m.vol, m.tmu, m.tlam, mu, lambda, partial);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_sum_fixed(nb, partial, out, s, partial + energy_partials(m));
}

hipError_t launch_inertia_energy(const DevMesh& m, double sdt, const ContactArgs& ct, double* partial, double* out,
                                 hipStream_t s) {
    const int nb = (m.nFree + 255) / 256;
    if (nb == 0) return launch_sum_fixed(0, partial, out, s);
    hipLaunchKernelGGL(k_inertia_energy, dim3(nb), dim3(256), 0, s, m.nFree, m.pos, m.inert, m.mass, sdt, ct, partial);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_sum_fixed(nb, partial, out, s, partial + energy_partials(m));
}

// ---------------------------------------------------------------------------------------
// Mesh build
// ---------------------------------------------------------------------------------------
constexpr int SCAN_B = 256, SCAN_I = 8, SCAN_CHUNK = SCAN_B * SCAN_I;

// exclusive scan of one chunk per block; chunk totals to sums[]
__global__ __launch_bounds__(SCAN_B) void k_scan_chunk(const int* __restrict__ in, int* __restrict__ out,
                                                       int* __restrict__ sums, int n) {
    __shared__ int sh[SCAN_B];
    const int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_I;
    int v[SCAN_I];
    int tot = 0;
    for (int i = 0; i < SCAN_I; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        tot += v[i];
    }
    sh[threadIdx.x] = tot;
    __syncthreads();
    for (int off = 1; off < SCAN_B; off <<= 1) {
        const int t = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    int ex = sh[threadIdx.x] - tot;
    for (int i = 0; i < SCAN_I; ++i) {
        if (base + i < n) out[base + i] = ex;
        ex += v[i];
    }
    if (threadIdx.x == SCAN_B - 1) sums[blockIdx.x] = sh[SCAN_B - 1];
}

__global__ __launch_bounds__(SCAN_B) void k_scan_add(int* __restrict__ out, const int* __restrict__ offs, int n) {
    const int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_I;
    const int o = offs[blockIdx.x];
    for (int i = 0; i < SCAN_I; ++i)
        if (base + i < n) out[base + i] += o;
}

__global__ void k_scan_total(const int* __restrict__ in, int* __restrict__ out, int n) {
    out[n] = n > 0 ? out[n - 1] + in[n - 1] : 0;
}

// out[0..n] = exclusive scan of in[0..n), out[n] = total.  out must hold n+1 ints.
static hipError_t scan_exclusive(const int* in, int* out, int n, hipStream_t s) {
    hipError_t e;
    if (n > 0) {
        const int nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
        int* sums = nullptr;
        int* soff = nullptr;
        if ((e = hipMalloc(&sums, sizeof(int) * nb)) != hipSuccess) return e;
        if ((e = hipMalloc(&soff, sizeof(int) * (nb + 1))) != hipSuccess) { (void)hipFree(sums); return e; }
        hipLaunchKernelGGL(k_scan_chunk, dim3(nb), dim3(SCAN_B), 0, s, in, out, sums, n);
        e = hipGetLastError();
        if (e == hipSuccess && nb > 1) {
            e = scan_exclusive(sums, soff, nb, s);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(SCAN_B), 0, s, out, soff, n);
                e = hipGetLastError();
            }
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(sums);
        (void)hipFree(soff);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(1), 0, s, in, out, n);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_degree(int nT, const i4* __restrict__ tets, int* __restrict__ deg) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const i4 tt = tets[t];
    atomicAdd(&deg[tt.x], 1);
    atomicAdd(&deg[tt.y], 1);
    atomicAdd(&deg[tt.z], 1);
    atomicAdd(&deg[tt.w], 1);
}

__global__ __launch_bounds__(256) void k_fill_adj(int nT, const i4* __restrict__ tets, const int* __restrict__ off,
                                                  int* __restrict__ cursor, int* __restrict__ entry) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const i4 tt = tets[t];
    const int v[4] = {tt.x, tt.y, tt.z, tt.w};
    for (int j = 0; j < 4; ++j) {
        const int slot = off[v[j]] + atomicAdd(&cursor[v[j]], 1);
        entry[slot] = t * 4 + j;
    }
}

// Atomic fill leaves each vertex's entries in arbitrary order; sorting by tet*4+corner
// restores ascending tet id, the order TetMesh::initialize appends them (TetMesh.h:254-266).
__global__ __launch_bounds__(256) void k_sort_segments(int nV, const int* __restrict__ off, int* __restrict__ entry) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    const int b = off[v], e = off[v + 1];
    for (int i = b + 1; i < e; ++i) {
        const int key = entry[i];
        int j = i - 1;
        while (j >= b && entry[j] > key) {
            entry[j + 1] = entry[j];
            --j;
        }
        entry[j + 1] = key;
    }
}

__global__ __launch_bounds__(256) void k_key_flag(int n, const int* __restrict__ key, int k, int* __restrict__ flag) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) flag[v] = key[v] == k ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_key_place(int n, const int* __restrict__ key, int k,
                                                   const int* __restrict__ rank, int base, int* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n && key[v] == k) out[base + rank[v]] = v;
}

__global__ __launch_bounds__(256) void k_slot_key(int n, const int* __restrict__ colorOf,
                                                  const uint8_t* __restrict__ fixed, int nColors,
                                                  int* __restrict__ key) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) key[v] = fixed[v] ? nColors : colorOf[v];
}

__global__ __launch_bounds__(256) void k_invert(int n, const int* __restrict__ perm, int* __restrict__ inv) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) inv[perm[s]] = s;
}

// Per-tet corner ids -> device slots (the per-tet kernels then gather positions directly).
__global__ __launch_bounds__(256) void k_tets_to_slots(int nT, const i4* __restrict__ tets, const int* __restrict__ inv,
                                                       i4* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const i4 c = tets[t];
    out[t] = i4{inv[c.x], inv[c.y], inv[c.z], inv[c.w]};
}

__global__ __launch_bounds__(256) void k_slot_valence(int nFree, const int* __restrict__ perm,
                                                      const int* __restrict__ off, int* __restrict__ cnt) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nFree) cnt[s] = off[perm[s] + 1] - off[perm[s]];
}

__global__ __launch_bounds__(256) void k_build_pairs(int nFree, const int* __restrict__ perm,
                                                     const int* __restrict__ inv, const int* __restrict__ off,
                                                     const int* __restrict__ entry, const i4* __restrict__ tets,
                                                     const float* __restrict__ dminv, const float* __restrict__ vol,
                                                     const float* __restrict__ tmu, const float* __restrict__ tlam,
                                                     const int* __restrict__ row, const int* __restrict__ chunkBase,
                                                     i4* __restrict__ ptet, f4* __restrict__ pa, f4* __restrict__ pb,
                                                     float* __restrict__ pc, f2* __restrict__ pmat,
                                                     f4* __restrict__ pq) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nFree) return;
    const int v = perm[s];
    int p = row[s];
    for (int e = off[v]; e < off[v + 1]; ++e, ++p) {
        const int t = entry[e] >> 2;
        const int k = entry[e] & 3;
        const i4 tt = tets[t];
        const int c[4] = {tt.x, tt.y, tt.z, tt.w};
        const float* D = dminv + 9 * (size_t)t;
        // record layout: see the pair contribution comment above (base corner 0 as in Ds)
        int o[3], rows[3];
        if (k == 0) {
            o[0] = inv[c[1]]; o[1] = inv[c[2]]; o[2] = inv[c[3]];
            rows[0] = 0; rows[1] = 1; rows[2] = 2;
        } else {
            o[0] = inv[c[0]];
            rows[0] = k - 1;
            int n = 1;
            for (int j = 1; j < 4; ++j)
                if (j != k) {
                    o[n] = inv[c[j]];
                    rows[n] = j - 1;
                    ++n;
                }
        }
        float Q[9];
        for (int n = 0; n < 3; ++n)
            for (int j = 0; j < 3; ++j) Q[3 * n + j] = D[3 * rows[n] + j];
        // owner offset from the chunk's first owner (checked on the host to fit 12 bits)
        const int own = (s - chunkBase[p >> CHUNK_SHIFT]) & 0xFFF;
        i4 r;
        r.x = o[0] | ((own & 15) << 27) | (k == 0 ? (int)0x80000000 : 0);
        r.y = o[1] | (((own >> 4) & 15) << 27);
        r.z = o[2] | (((own >> 8) & 15) << 27);
        r.w = __float_as_int(vol[t]);
        ptet[p] = r;
        if (pa) {  // full records (built here only when they are resident already)
            f4 a, b;
            a.x = Q[0]; a.y = Q[1]; a.z = Q[2]; a.w = Q[3];
            b.x = Q[4]; b.y = Q[5]; b.z = Q[6]; b.w = Q[7];
            pa[p] = a;
            pb[p] = b;
            pc[p] = Q[8];
        }
        {
            // compact record (tiled kernel), in double from the FP32 DmInv rows R_n:
            // q = -(R0+R1+R2) (corner 0) or R0, alpha_n = R_n . q, beta = -alpha_0 (corner 0) or
            // alpha_0, d = det(DmInv) * (-1, +1, -1, +1)[k]
            double R[3][3], q[3], al[3];
            for (int n = 0; n < 3; ++n)
                for (int j = 0; j < 3; ++j) R[n][j] = (double)Q[3 * n + j];
            for (int j = 0; j < 3; ++j) q[j] = k == 0 ? -(R[0][j] + R[1][j] + R[2][j]) : R[0][j];
            for (int n = 0; n < 3; ++n) al[n] = R[n][0] * q[0] + R[n][1] * q[1] + R[n][2] * q[2];
            double M[9];
            for (int j = 0; j < 9; ++j) M[j] = (double)D[j];
            const double det = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
                               M[2] * (M[3] * M[7] - M[4] * M[6]);
            f4 cq;
            cq.x = (float)(k == 0 ? -al[0] : al[0]);
            cq.y = (float)al[1];
            cq.z = (float)al[2];
            cq.w = (float)((k == 0 || k == 2) ? -det : det);
            pq[p] = cq;
        }
        if (pmat) {
            f2 mm;
            mm.x = tmu[t];
            mm.y = tlam[t];
            pmat[p] = mm;
        }
    }
}

// The full records (three DmInv rows per record, as k_build_pairs orders them) of every free
// slot, from the per-tet arrays kept on the device (corner slots in caller tet order, DmInv
// row-major): built on first use by the kernels that read them.
__global__ __launch_bounds__(256) void k_build_full(int nFree, const int* __restrict__ perm,
                                                    const int* __restrict__ off, const int* __restrict__ entry,
                                                    const float* __restrict__ dminv, const int* __restrict__ row,
                                                    f4* __restrict__ pa, f4* __restrict__ pb, float* __restrict__ pc) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nFree) return;
    const int v = perm[s];
    int p = row[s];
    for (int e = off[v]; e < off[v + 1]; ++e, ++p) {
        const int t = entry[e] >> 2;
        const int k = entry[e] & 3;
        const float* D = dminv + 9 * (size_t)t;
        // rows: k == 0 -> 0, 1, 2; else k-1, then the other two in tet order
        int rows[3];
        if (k == 0) {
            rows[0] = 0; rows[1] = 1; rows[2] = 2;
        } else {
            rows[0] = k - 1;
            int n = 1;
            for (int j = 1; j < 4; ++j)
                if (j != k) rows[n++] = j - 1;
        }
        f4 a, b;
        a.x = D[3 * rows[0]]; a.y = D[3 * rows[0] + 1]; a.z = D[3 * rows[0] + 2]; a.w = D[3 * rows[1]];
        b.x = D[3 * rows[1] + 1]; b.y = D[3 * rows[1] + 2]; b.z = D[3 * rows[2]]; b.w = D[3 * rows[2] + 1];
        pa[p] = a;
        pb[p] = b;
        pc[p] = D[3 * rows[2] + 2];
    }
}

hipError_t ensure_full_records(DevMesh& m, hipStream_t s, std::string& err) {
    if (m.pa) return hipSuccess;
    const size_t n = (size_t)std::max<int64_t>(m.nPairs, 1);
    hipError_t e = hipMalloc(&m.pa, sizeof(f4) * n);
    if (e == hipSuccess) e = hipMalloc(&m.pb, sizeof(f4) * n);
    if (e == hipSuccess) e = hipMalloc(&m.pc, sizeof(float) * n);
    if (e == hipSuccess && m.nFree > 0) {
        hipLaunchKernelGGL(k_build_full, dim3((unsigned)((m.nFree + 255) / 256)), dim3(256), 0, s, m.nFree, m.perm,
                           m.adjOff, m.adjEntry, m.dminv, m.row, m.pa, m.pb, m.pc);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        err = std::string("full pair records: ") + hipGetErrorString(e);
        void* ps[] = {m.pa, m.pb, m.pc};
        for (void* q : ps)
            if (q) (void)hipFree(q);
        m.pa = m.pb = nullptr;
        m.pc = nullptr;
        return e;
    }
    m.bytes += (int64_t)n * (int64_t)(2 * sizeof(f4) + sizeof(float));
    return hipSuccess;
}

void release_full_records(DevMesh& m, hipStream_t s) {
    (void)hipStreamSynchronize(s);
    if (m.pa) {
        const size_t n = (size_t)std::max<int64_t>(m.nPairs, 1);
        void* ps[] = {m.pa, m.pb, m.pc};
        for (void* q : ps)
            if (q) (void)hipFree(q);
        m.pa = m.pb = nullptr;
        m.pc = nullptr;
        m.bytes -= (int64_t)n * (int64_t)(2 * sizeof(f4) + sizeof(float));
    }
    if (m.tpa) {  // the tile cut's padded copies of them
        const size_t n = (size_t)std::max(m.tileEntries, 1) * (size_t)m.tileRecords;
        void* ps[] = {m.tpa, m.tpb, m.tpc};
        for (void* q : ps)
            if (q) (void)hipFree(q);
        m.tpa = m.tpb = nullptr;
        m.tpc = nullptr;
        const int64_t b = (int64_t)n * (int64_t)(2 * sizeof(f4) + sizeof(float));
        m.padBytes -= b;
        m.bytes -= b;
    }
}

__global__ __launch_bounds__(256) void k_gather_scalar(int n, const int* __restrict__ perm,
                                                       const float* __restrict__ src, float* __restrict__ dst) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) dst[s] = src[perm[s]];
}

template <class T>
static hipError_t dalloc(DevMesh& m, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
    if (e == hipSuccess) m.bytes += (int64_t)(sizeof(T) * count);
    return e;
}

static inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

#define VBD_TRY(expr)                                   \
    do {                                                \
        hipError_t _e = (expr);                         \
        if (_e != hipSuccess) {                         \
            rc = _e;                                    \
            err = std::string(#expr) + ": " + hipGetErrorString(_e); \
            goto fail;                                  \
        }                                               \
    } while (0)

hipError_t build_device_mesh(DevMesh& m, int nV, int nT, int nColors, const i4* d_tets, const int* d_colorOf,
                             const uint8_t* d_fixed, const float* d_dminv, const float* d_vol, const float* d_mu,
                             const float* d_lam, const float* d_x3, const float* d_y3, const float* d_v3,
                             const float* d_mass, const int* h_perm, const int* h_slot_off, hipStream_t s,
                             std::string& err) {
    hipError_t rc = hipSuccess;
    int* deg = nullptr;
    int* cursor = nullptr;
    int* key = nullptr;
    int* flag = nullptr;
    int* rank = nullptr;
    int* cnt = nullptr;
    std::vector<int> hrow_last(1, 0);
    m.nV = nV;
    m.nT = nT;
    m.nColors = nColors;
    m.perTetMat = d_mu != nullptr;
    m.colorSlotOff.assign(nColors + 1, 0);
    m.colorFullOff.assign(nColors + 1, 0);

    // ---- vertex -> tet CSR in ascending tet order ----
    VBD_TRY(dalloc(m, &m.adjOff, (size_t)nV + 1));
    VBD_TRY(dalloc(m, &m.adjEntry, (size_t)4 * nT));
    VBD_TRY(hipMalloc(&deg, sizeof(int) * (nV + 1)));
    VBD_TRY(hipMalloc(&cursor, sizeof(int) * (nV + 1)));
    VBD_TRY(hipMemsetAsync(deg, 0, sizeof(int) * (nV + 1), s));
    VBD_TRY(hipMemsetAsync(cursor, 0, sizeof(int) * (nV + 1), s));
    if (nT > 0) {
        hipLaunchKernelGGL(k_degree, grid_for(nT), dim3(256), 0, s, nT, d_tets, deg);
        VBD_TRY(hipGetLastError());
    }
    VBD_TRY(scan_exclusive(deg, m.adjOff, nV, s));
    if (nT > 0) {
        hipLaunchKernelGGL(k_fill_adj, grid_for(nT), dim3(256), 0, s, nT, d_tets, m.adjOff, cursor, m.adjEntry);
        VBD_TRY(hipGetLastError());
    }
    if (nV > 0) {
        hipLaunchKernelGGL(k_sort_segments, grid_for(nV), dim3(256), 0, s, nV, m.adjOff, m.adjEntry);
        VBD_TRY(hipGetLastError());
    }

    // ---- per-color lists (every vertex) and color-major slot order (free vertices) ----
    VBD_TRY(dalloc(m, &m.colorList, (size_t)nV));
    VBD_TRY(dalloc(m, &m.perm, (size_t)nV));
    VBD_TRY(dalloc(m, &m.inv, (size_t)nV));
    VBD_TRY(hipMalloc(&key, sizeof(int) * (nV + 1)));
    VBD_TRY(hipMalloc(&flag, sizeof(int) * (nV + 1)));
    VBD_TRY(hipMalloc(&rank, sizeof(int) * (nV + 1)));
    {
        int base = 0;
        for (int c = 0; c < nColors; ++c) {
            int total = 0;
            if (nV > 0) {
                hipLaunchKernelGGL(k_key_flag, grid_for(nV), dim3(256), 0, s, nV, d_colorOf, c, flag);
                VBD_TRY(hipGetLastError());
            }
            VBD_TRY(scan_exclusive(flag, rank, nV, s));
            VBD_TRY(hipMemcpyAsync(&total, rank + nV, sizeof(int), hipMemcpyDeviceToHost, s));
            VBD_TRY(hipStreamSynchronize(s));
            if (total > 0) {
                hipLaunchKernelGGL(k_key_place, grid_for(nV), dim3(256), 0, s, nV, d_colorOf, c, rank, base,
                                   m.colorList);
                VBD_TRY(hipGetLastError());
            }
            base += total;
            m.colorFullOff[c + 1] = base;
        }
        if (base != nV) {
            err = "color lists do not cover every vertex exactly once";
            rc = hipErrorInvalidValue;
            goto fail;
        }
        if (h_perm) {
            // slot order chosen by the host (color-major, spatial order inside a color)
            VBD_TRY(hipMemcpyAsync(m.perm, h_perm, sizeof(int) * (size_t)nV, hipMemcpyHostToDevice, s));
            for (int c = 0; c <= nColors; ++c) m.colorSlotOff[c] = h_slot_off[c];
        } else {
            if (nV > 0) {
                hipLaunchKernelGGL(k_slot_key, grid_for(nV), dim3(256), 0, s, nV, d_colorOf, d_fixed, nColors, key);
                VBD_TRY(hipGetLastError());
            }
            base = 0;
            for (int c = 0; c <= nColors; ++c) {
                int total = 0;
                if (nV > 0) {
                    hipLaunchKernelGGL(k_key_flag, grid_for(nV), dim3(256), 0, s, nV, key, c, flag);
                    VBD_TRY(hipGetLastError());
                }
                VBD_TRY(scan_exclusive(flag, rank, nV, s));
                VBD_TRY(hipMemcpyAsync(&total, rank + nV, sizeof(int), hipMemcpyDeviceToHost, s));
                VBD_TRY(hipStreamSynchronize(s));
                if (total > 0) {
                    hipLaunchKernelGGL(k_key_place, grid_for(nV), dim3(256), 0, s, nV, key, c, rank, base, m.perm);
                    VBD_TRY(hipGetLastError());
                }
                base += total;
                if (c < nColors) m.colorSlotOff[c + 1] = base;
            }
        }
        m.nFree = m.colorSlotOff[nColors];
        if (nV > 0) {
            hipLaunchKernelGGL(k_invert, grid_for(nV), dim3(256), 0, s, nV, m.perm, m.inv);
            VBD_TRY(hipGetLastError());
        }
    }

    // ---- pair records in slot order ----
    VBD_TRY(dalloc(m, &m.row, (size_t)m.nFree + 1));
    VBD_TRY(hipMalloc(&cnt, sizeof(int) * (m.nFree + 1)));
    if (m.nFree > 0) {
        hipLaunchKernelGGL(k_slot_valence, grid_for(m.nFree), dim3(256), 0, s, m.nFree, m.perm, m.adjOff, cnt);
        VBD_TRY(hipGetLastError());
    }
    VBD_TRY(scan_exclusive(cnt, m.row, m.nFree, s));
    {
        std::vector<int> hcnt(m.nFree > 0 ? m.nFree : 1, 0);
        if (m.nFree > 0) VBD_TRY(hipMemcpyAsync(hcnt.data(), cnt, sizeof(int) * m.nFree, hipMemcpyDeviceToHost, s));
        VBD_TRY(hipMemcpyAsync(hrow_last.data(), m.row + m.nFree, sizeof(int), hipMemcpyDeviceToHost, s));
        VBD_TRY(hipStreamSynchronize(s));
        m.nPairs = hrow_last[0];
        m.maxValence = 0;
        m.valenceHist.assign(1, 0);
        m.hostRow.assign((size_t)m.nFree + 1, 0);
        for (int i = 0; i < m.nFree; ++i) {
            const int d = hcnt[i];
            if (d > m.maxValence) m.maxValence = d;
            if ((int)m.valenceHist.size() <= d) m.valenceHist.resize(d + 1, 0);
            m.valenceHist[d]++;
            m.hostRow[i + 1] = m.hostRow[i] + d;
        }
        // chunk c's base = owner slot of record c * 4096; every record's owner must lie within
        // 4095 slots of its chunk's base (fails only with long runs of tet-less free vertices).
        // One spare entry: a tile of tet-less vertices after the last record reads the base
        // of chunk nPairs >> 12.
        const int64_t nChunks = (m.nPairs + (1 << CHUNK_SHIFT) - 1) >> CHUNK_SHIFT;
        std::vector<int> hcb((size_t)nChunks + 1, 0);
        m.tiledOk = true;
        for (int i = 0; i < m.nFree; ++i) {
            const int a = m.hostRow[i], b = m.hostRow[i + 1];
            for (int64_t c = ((int64_t)a + (1 << CHUNK_SHIFT) - 1) >> CHUNK_SHIFT; (c << CHUNK_SHIFT) < b; ++c)
                hcb[c] = i;
            if (a < b && i - hcb[a >> CHUNK_SHIFT] > 4095) m.tiledOk = false;
        }
        VBD_TRY(dalloc(m, &m.chunkBase, hcb.size()));
        VBD_TRY(hipMemcpyAsync(m.chunkBase, hcb.data(), sizeof(int) * hcb.size(), hipMemcpyHostToDevice, s));
        VBD_TRY(hipStreamSynchronize(s));
    }
    VBD_TRY(dalloc(m, &m.ptet, (size_t)m.nPairs));
    // the full records (pa, pb, pc) are built on first use (ensure_full_records)
    if (m.perTetMat) VBD_TRY(dalloc(m, &m.pmat, (size_t)m.nPairs));
    VBD_TRY(dalloc(m, &m.pq, (size_t)m.nPairs));
    if (m.nFree > 0) {
        hipLaunchKernelGGL(k_build_pairs, grid_for(m.nFree), dim3(256), 0, s, m.nFree, m.perm, m.inv, m.adjOff,
                           m.adjEntry, d_tets, d_dminv, d_vol, d_mu, d_lam, m.row, m.chunkBase, m.ptet, m.pa,
                           m.pb, m.pc, m.perTetMat ? m.pmat : nullptr, m.pq);
        VBD_TRY(hipGetLastError());
    }

    // ---- per-tet arrays kept for the energy reduction ----
    VBD_TRY(dalloc(m, &m.tets, (size_t)nT));
    VBD_TRY(dalloc(m, &m.dminv, (size_t)9 * nT));
    VBD_TRY(dalloc(m, &m.vol, (size_t)nT));
    if (nT > 0) {
        hipLaunchKernelGGL(k_tets_to_slots, grid_for(nT), dim3(256), 0, s, nT, d_tets, m.inv, m.tets);
        VBD_TRY(hipGetLastError());
        VBD_TRY(hipMemcpyAsync(m.dminv, d_dminv, sizeof(float) * 9 * nT, hipMemcpyDeviceToDevice, s));
        VBD_TRY(hipMemcpyAsync(m.vol, d_vol, sizeof(float) * nT, hipMemcpyDeviceToDevice, s));
    }
    if (m.perTetMat) {
        VBD_TRY(dalloc(m, &m.tmu, (size_t)nT));
        VBD_TRY(dalloc(m, &m.tlam, (size_t)nT));
        if (nT > 0) {
            VBD_TRY(hipMemcpyAsync(m.tmu, d_mu, sizeof(float) * nT, hipMemcpyDeviceToDevice, s));
            VBD_TRY(hipMemcpyAsync(m.tlam, d_lam, sizeof(float) * nT, hipMemcpyDeviceToDevice, s));
        }
    }

    // ---- vertex state in slot order ----
    VBD_TRY(dalloc(m, &m.pos, (size_t)nV));
    VBD_TRY(dalloc(m, &m.inert, (size_t)nV));
    VBD_TRY(dalloc(m, &m.vel, (size_t)nV));
    VBD_TRY(dalloc(m, &m.xprev, (size_t)nV));
    VBD_TRY(dalloc(m, &m.save[0], (size_t)nV));
    VBD_TRY(dalloc(m, &m.save[1], (size_t)nV));
    VBD_TRY(dalloc(m, &m.mass, (size_t)nV));
    VBD_TRY(hipMemsetAsync(m.save[0], 0, sizeof(f4) * (nV > 0 ? nV : 1), s));
    VBD_TRY(hipMemsetAsync(m.save[1], 0, sizeof(f4) * (nV > 0 ? nV : 1), s));
    VBD_TRY(hipMemsetAsync(m.xprev, 0, sizeof(f4) * (nV > 0 ? nV : 1), s));
    VBD_TRY(launch_scatter_pos(nV, d_x3, m.inv, m.pos, s));
    VBD_TRY(launch_scatter_pos(nV, d_y3 ? d_y3 : d_x3, m.inv, m.inert, s));
    if (d_v3) {
        VBD_TRY(launch_scatter_pos(nV, d_v3, m.inv, m.vel, s));
    } else {
        VBD_TRY(hipMemsetAsync(m.vel, 0, sizeof(f4) * (nV > 0 ? nV : 1), s));
    }
    if (nV > 0) {
        hipLaunchKernelGGL(k_gather_scalar, grid_for(nV), dim3(256), 0, s, nV, m.perm, d_mass, m.mass);
        VBD_TRY(hipGetLastError());
    }
    VBD_TRY(hipStreamSynchronize(s));
    (void)hipFree(deg);
    (void)hipFree(cursor);
    (void)hipFree(key);
    (void)hipFree(flag);
    (void)hipFree(rank);
    (void)hipFree(cnt);
    return hipSuccess;
fail:
    (void)hipStreamSynchronize(s);
    if (deg) (void)hipFree(deg);
    if (cursor) (void)hipFree(cursor);
    if (key) (void)hipFree(key);
    if (flag) (void)hipFree(flag);
    if (rank) (void)hipFree(rank);
    if (cnt) (void)hipFree(cnt);
    free_device_mesh(m);
    return rc == hipSuccess ? hipErrorUnknown : rc;
}
#undef VBD_TRY

void free_device_mesh(DevMesh& m) {
    void* ptrs[] = {m.pos, m.inert, m.vel, m.xprev, m.save[0], m.save[1], m.mass, m.row, m.perm, m.inv,
                    m.ptet, m.pa, m.pb, m.pc, m.pmat, m.adjOff, m.adjEntry, m.colorList, m.tets, m.dminv,
                    m.vol, m.tmu, m.tlam, m.chunkBase, m.tileStart, m.tileRec, m.pq,
                    m.tpt, m.tpq, m.tpm, m.tpa, m.tpb, m.tpc, m.dColorTileBase};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    m = DevMesh();
}

}  // namespace vbd
