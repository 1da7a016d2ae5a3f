// FP64 device path of the differentiation entry points (the converged step, its residual, the
// Hessian-vector product, the adjoint solve and the parameter gradients).
//
// The FP32 state in absolute coordinates cannot converge an implicit step further than about
// H ulp(x) per vertex (the local solve's update rounds away below that), so the FP32 converged
// step stalls near 5e-3 |M g| on the 16^3 workload.  With the FP64 state the same Gauss-Seidel
// iteration converges to |grad G| < 1e-6.  The kernels read the full pair records (the FP32
// DmInv rows, promoted exactly) and compute everything else in double:
//   F = Ds DmInv with Ds's edges from corner 0 (Energy.cpp:11-21), C = cof F (Energy.cpp:61-73),
//   P = mu (F - C) + lambda (J - 1) C, g = V P q, H = V [mu |q|^2 I + lambda (C q)(C q)^T]
//   (Energy.cpp:90-158 in closed form, the vertex's diagonal block), the inertia term
//   m/h^2 (VBD.cpp:170-173), the penalty contact, and the adjugate 3x3 solve skipped when
//   |det H| <= 1e-8 (VBD.cpp:141-153, Types.h:85-112); the HVP's cofactor derivative follows
//   Energy.cpp:76-88.
// Vertex vectors are double4 per slot (w unused), in the same slot order as the FP32 arrays.
#include "vbd_internal.h"

namespace vbd {

// lanes per vertex: 32 covers a tet-mesh vertex's records (23 on average on a Kuhn grid) in one
// step, which shortens each launch's dependent chain; on the 16^3 differentiable cube (launch-
// bound, about 600 vertices per color) forward and adjoint ran 21-24% faster than with 8 lanes,
// 11-13% faster than with 16 and 4% faster than with 64
constexpr int LPV64 = 32;

constexpr int SLOT_MASK64 = 0x07FFFFFF;  // the record's corner slots (the owner bits above are unused here)

struct Blk64 {
    double h00, h01, h02, h11, h12, h22, r0, r1, r2;
};

struct Kin64 {
    double F[3][3], C[3][3], J, q[3];
};

__device__ __forceinline__ Kin64 kin64(const bool corner0, const f4 a, const f4 b, const float c, const d4 xs,
                                       const d4 ya, const d4 yb, const d4 yc) {
    const double R[3][3] = {{a.x, a.y, a.z}, {a.w, b.x, b.y}, {b.z, b.w, c}};
    const d4 base = corner0 ? xs : ya;
    const d4 first = corner0 ? ya : xs;
    const double e[3][3] = {{first.x - base.x, yb.x - base.x, yc.x - base.x},
                            {first.y - base.y, yb.y - base.y, yc.y - base.y},
                            {first.z - base.z, yb.z - base.z, yc.z - base.z}};
    Kin64 k;
    for (int r = 0; r < 3; ++r)
        for (int j = 0; j < 3; ++j) k.F[r][j] = e[r][0] * R[0][j] + e[r][1] * R[1][j] + e[r][2] * R[2][j];
    const double(&F)[3][3] = k.F;
    k.C[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
    k.C[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
    k.C[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
    k.C[1][0] = F[2][1] * F[0][2] - F[2][2] * F[0][1];
    k.C[1][1] = F[2][2] * F[0][0] - F[2][0] * F[0][2];
    k.C[1][2] = F[2][0] * F[0][1] - F[2][1] * F[0][0];
    k.C[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
    k.C[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
    k.C[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    k.J = F[0][0] * k.C[0][0] + F[0][1] * k.C[0][1] + F[0][2] * k.C[0][2];
    for (int j = 0; j < 3; ++j) k.q[j] = corner0 ? -(R[0][j] + R[1][j] + R[2][j]) : R[0][j];
    return k;
}

// g and the diagonal block of one (vertex, tet) record; f -= g, H += block
__device__ __forceinline__ void accumulate64(const Kin64& k, const double V, const double mu, const double lam,
                                             Blk64& A) {
    double Cq[3], Fq[3];
    for (int r = 0; r < 3; ++r) {
        Cq[r] = k.C[r][0] * k.q[0] + k.C[r][1] * k.q[1] + k.C[r][2] * k.q[2];
        Fq[r] = k.F[r][0] * k.q[0] + k.F[r][1] * k.q[1] + k.F[r][2] * k.q[2];
    }
    const double aj = lam * (k.J - 1.0);
    const double vm = V * mu * (k.q[0] * k.q[0] + k.q[1] * k.q[1] + k.q[2] * k.q[2]);
    const double vl = V * lam;
    A.h00 += vm + vl * Cq[0] * Cq[0];
    A.h01 += vl * Cq[0] * Cq[1];
    A.h02 += vl * Cq[0] * Cq[2];
    A.h11 += vm + vl * Cq[1] * Cq[1];
    A.h12 += vl * Cq[1] * Cq[2];
    A.h22 += vm + vl * Cq[2] * Cq[2];
    A.r0 -= V * (mu * (Fq[0] - Cq[0]) + aj * Cq[0]);
    A.r1 -= V * (mu * (Fq[1] - Cq[1]) + aj * Cq[1]);
    A.r2 -= V * (mu * (Fq[2] - Cq[2]) + aj * Cq[2]);
}

// (K_t w)_s = V dP q with dF = the w-differences through DmInv (Energy.cpp:76-88, :123-127)
__device__ __forceinline__ void hvp64(const Kin64& k, const bool corner0, const f4 a, const f4 b, const float c,
                                      const double V, const double mu, const double lam, const d4 ws, const d4 wa,
                                      const d4 wb, const d4 wc, double o[3]) {
    const double R[3][3] = {{a.x, a.y, a.z}, {a.w, b.x, b.y}, {b.z, b.w, c}};
    const d4 base = corner0 ? ws : wa;
    const d4 first = corner0 ? wa : ws;
    const double e[3][3] = {{first.x - base.x, wb.x - base.x, wc.x - base.x},
                            {first.y - base.y, wb.y - base.y, wc.y - base.y},
                            {first.z - base.z, wb.z - base.z, wc.z - base.z}};
    double G[3][3];
    for (int r = 0; r < 3; ++r)
        for (int j = 0; j < 3; ++j) G[r][j] = e[r][0] * R[0][j] + e[r][1] * R[1][j] + e[r][2] * R[2][j];
    const double(&F)[3][3] = k.F;
    double dC[3][3];
    dC[0][0] = G[1][1] * F[2][2] + F[1][1] * G[2][2] - G[1][2] * F[2][1] - F[1][2] * G[2][1];
    dC[0][1] = G[1][2] * F[2][0] + F[1][2] * G[2][0] - G[1][0] * F[2][2] - F[1][0] * G[2][2];
    dC[0][2] = G[1][0] * F[2][1] + F[1][0] * G[2][1] - G[1][1] * F[2][0] - F[1][1] * G[2][0];
    dC[1][0] = G[2][1] * F[0][2] + F[2][1] * G[0][2] - G[2][2] * F[0][1] - F[2][2] * G[0][1];
    dC[1][1] = G[2][2] * F[0][0] + F[2][2] * G[0][0] - G[2][0] * F[0][2] - F[2][0] * G[0][2];
    dC[1][2] = G[2][0] * F[0][1] + F[2][0] * G[0][1] - G[2][1] * F[0][0] - F[2][1] * G[0][0];
    dC[2][0] = G[0][1] * F[1][2] + F[0][1] * G[1][2] - G[0][2] * F[1][1] - F[0][2] * G[1][1];
    dC[2][1] = G[0][2] * F[1][0] + F[0][2] * G[1][0] - G[0][0] * F[1][2] - F[0][0] * G[1][2];
    dC[2][2] = G[0][0] * F[1][1] + F[0][0] * G[1][1] - G[0][1] * F[1][0] - F[0][1] * G[1][0];
    double dJ = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int j = 0; j < 3; ++j) dJ += k.C[r][j] * G[r][j];
    const double ld = lam * dJ, aj = lam * (k.J - 1.0);
    for (int r = 0; r < 3; ++r) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j) s += (mu * (G[r][j] - dC[r][j]) + ld * k.C[r][j] + aj * dC[r][j]) * k.q[j];
        o[r] = V * s;
    }
}

// Sum over the aligned group of LPV64 = 32 lanes, valid in the group's last lane (the only lane
// that reads it): DPP suffix sums inside each row of 16 (row_shr 1, 2, 4, 8), then row 0's total
// carried into row 1 and row 2's into row 3 (row_bcast:15), the two 32-bit halves of a double
// moved separately.  Lanes whose DPP source lies outside the row receive 0.  A fixed order, and
// no LDS round trip per step as a cross-lane shuffle takes.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_take64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double gsum64(double v) {
    static_assert(LPV64 == 32, "the DPP tree sums groups of 32 lanes");
    v += dpp_take64<0x111, 0xF>(v);  // row_shr:1
    v += dpp_take64<0x112, 0xF>(v);  // row_shr:2
    v += dpp_take64<0x114, 0xF>(v);  // row_shr:4
    v += dpp_take64<0x118, 0xF>(v);  // row_shr:8
    v += dpp_take64<0x142, 0xA>(v);  // row_bcast:15 into rows 1, 3
    return v;
}

enum Pass64 {
    P64_FWD_UPDATE = 0, P64_FWD_RESID = 1, P64_ADJ_SOLVE = 2, P64_ADJ_RESID = 3, P64_HVP = 4, P64_DIAG_INV = 5
};

// One pass over free slots [slot0, slot0 + count), LPV64 lanes per vertex.
//   P64_FWD_UPDATE: one Gauss-Seidel vertex update of x (a color's launch); SAVE keeps x before it
//   P64_FWD_RESID:  out = grad G = m/h^2 (x - y) + sum g - contact force, r2 = |out|^2
//   P64_ADJ_SOLVE:  w_s += D_ss^-1 (rhs - A w)_s (a color's launch); SAVE keeps w before it
//   P64_ADJ_RESID:  out = rhs - A w, r2 = |out|^2
//   P64_HVP:        out = A w; with r2, r2 = w_s . (A w)_s
//   P64_DIAG_INV:   D_ss^-1 of the diagonal block (the ADJ_SOLVE block), symmetric, as
//                   out = (i00, i01, i02, i11), save = (i12, i22, 0, 0); zero when |det D| <= 1e-8
template <int MODE, bool MAT, bool CONTACT, bool SAVE>
__global__ __launch_bounds__(256) void k_pass64(int slot0, int count, d4* __restrict__ pos, const d4* __restrict__ inert,
                                                const float* __restrict__ mass, const int* __restrict__ row,
                                                const i4* __restrict__ ptet, const f4* __restrict__ pa,
                                                const f4* __restrict__ pb, const float* __restrict__ pc,
                                                const f2* __restrict__ pmat, d4* __restrict__ w,
                                                const d4* __restrict__ rhs, d4* __restrict__ out,
                                                double* __restrict__ r2, d4* __restrict__ save, double mu, double lambda,
                                                double sdt, ContactArgs ct) {
    constexpr bool NEED_W = MODE == P64_ADJ_SOLVE || MODE == P64_ADJ_RESID || MODE == P64_HVP;
    constexpr bool NEED_BLOCK = MODE == P64_FWD_UPDATE || MODE == P64_ADJ_SOLVE || MODE == P64_DIAG_INV;
    constexpr bool NEED_FORCE = MODE == P64_FWD_UPDATE || MODE == P64_FWD_RESID;
    const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x) / LPV64;
    const int lane = (int)threadIdx.x % LPV64;
    const bool valid = g < count;
    const int s = slot0 + (valid ? g : 0);
    Blk64 A = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double kw[3] = {0.0, 0.0, 0.0};
    d4 xs = {0, 0, 0, 0}, ws = {0, 0, 0, 0};
    if (valid) {
        xs = pos[s];
        if constexpr (NEED_W) ws = w[s];
        const int p1 = row[s + 1];
        for (int p = row[s] + lane; p < p1; p += LPV64) {
            const i4 t = ptet[p];
            const f4 a = pa[p], b = pb[p];
            const float c = pc[p];
            double tmu = mu, tlam = lambda;
            if constexpr (MAT) {
                const f2 mm = pmat[p];
                tmu = mm.x;
                tlam = mm.y;
            }
            const int sa = t.x & SLOT_MASK64, sb = t.y & SLOT_MASK64, sc = t.z & SLOT_MASK64;
            const bool c0 = t.x < 0;
            const double V = (double)__int_as_float(t.w);
            const Kin64 k = kin64(c0, a, b, c, xs, pos[sa], pos[sb], pos[sc]);
            if constexpr (NEED_FORCE || MODE == P64_DIAG_INV) {
                accumulate64(k, V, tmu, tlam, A);
            } else {
                if constexpr (MODE == P64_ADJ_SOLVE) {
                    Blk64 D = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                    accumulate64(k, V, tmu, tlam, D);
                    A.h00 += D.h00; A.h01 += D.h01; A.h02 += D.h02;
                    A.h11 += D.h11; A.h12 += D.h12; A.h22 += D.h22;
                }
                double o[3];
                hvp64(k, c0, a, b, c, V, tmu, tlam, ws, w[sa], w[sb], w[sc], o);
                kw[0] += o[0];
                kw[1] += o[1];
                kw[2] += o[2];
            }
        }
    }
    double h00 = 0, h01 = 0, h02 = 0, h11 = 0, h12 = 0, h22 = 0;
    if constexpr (NEED_BLOCK) {
        h00 = gsum64(A.h00); h01 = gsum64(A.h01); h02 = gsum64(A.h02);
        h11 = gsum64(A.h11); h12 = gsum64(A.h12); h22 = gsum64(A.h22);
    }
    double f0 = 0, f1 = 0, f2v = 0;
    if constexpr (NEED_FORCE) {
        f0 = gsum64(A.r0); f1 = gsum64(A.r1); f2v = gsum64(A.r2);
    } else if constexpr (MODE != P64_DIAG_INV) {
        kw[0] = gsum64(kw[0]); kw[1] = gsum64(kw[1]); kw[2] = gsum64(kw[2]);
    }
    if (!valid || lane != LPV64 - 1) return;
    const double m = mass[s];
    const double ms = m * sdt;
    // penalty contact: k (plane - x_a) e_a force and k e_a e_a^T Hessian while penetrating
    double ck = 0.0, cd = 0.0;
    int cax = 0;
    if constexpr (CONTACT) {
        cax = ct.axis;
        const double xa = cax == 0 ? xs.x : (cax == 1 ? xs.y : xs.z);
        cd = (double)ct.plane - xa;
        if (cd > 0.0) ck = ct.k;
        else cd = 0.0;
    }
    double e0 = 0, e1 = 0, e2 = 0;  // the right-hand side of the local system
    if constexpr (MODE == P64_DIAG_INV) {
        h00 += ms; h11 += ms; h22 += ms;
        if (cax == 0) h00 += ck;
        else if (cax == 1) h11 += ck;
        else h22 += ck;
        const double a00 = h11 * h22 - h12 * h12, a01 = h02 * h12 - h01 * h22, a02 = h01 * h12 - h02 * h11;
        const double a11 = h00 * h22 - h02 * h02, a12 = h02 * h01 - h00 * h12, a22 = h00 * h11 - h01 * h01;
        const double det = h00 * a00 + h01 * a01 + h02 * a02;
        const double inv = fabs(det) > 1e-8 ? 1.0 / det : 0.0;  // skipped blocks as in ADJ_SOLVE
        out[s] = d4{inv * a00, inv * a01, inv * a02, inv * a11};
        save[s] = d4{inv * a12, inv * a22, 0.0, 0.0};
        return;
    } else if constexpr (NEED_FORCE) {
        const d4 ys = inert[s];
        e0 = ms * (ys.x - xs.x) + f0;
        e1 = ms * (ys.y - xs.y) + f1;
        e2 = ms * (ys.z - xs.z) + f2v;
        if (cax == 0) e0 += ck * cd;
        else if (cax == 1) e1 += ck * cd;
        else e2 += ck * cd;
        if constexpr (MODE == P64_FWD_RESID) {
            d4 o = {-e0, -e1, -e2, 0.0};
            out[s] = o;
            r2[s] = e0 * e0 + e1 * e1 + e2 * e2;
            return;
        }
    } else {
        double aw[3] = {ms * ws.x + kw[0], ms * ws.y + kw[1], ms * ws.z + kw[2]};
        if (cax == 0) aw[0] += ck * ws.x;
        else if (cax == 1) aw[1] += ck * ws.y;
        else aw[2] += ck * ws.z;
        if constexpr (MODE == P64_HVP) {
            d4 o = {aw[0], aw[1], aw[2], 0.0};
            out[s] = o;
            if (r2) r2[s] = ws.x * aw[0] + ws.y * aw[1] + ws.z * aw[2];
            return;
        }
        const d4 b4 = rhs[s];
        e0 = b4.x - aw[0];
        e1 = b4.y - aw[1];
        e2 = b4.z - aw[2];
        if constexpr (MODE == P64_ADJ_RESID) {
            d4 o = {e0, e1, e2, 0.0};
            out[s] = o;
            r2[s] = e0 * e0 + e1 * e1 + e2 * e2;
            return;
        }
    }
    // P64_FWD_UPDATE / P64_ADJ_SOLVE: the 3x3 solve
    h00 += ms; h11 += ms; h22 += ms;
    if (cax == 0) h00 += ck;
    else if (cax == 1) h11 += ck;
    else h22 += ck;
    const double a00 = h11 * h22 - h12 * h12, a01 = h02 * h12 - h01 * h22, a02 = h01 * h12 - h02 * h11;
    const double a11 = h00 * h22 - h02 * h02, a12 = h02 * h01 - h00 * h12, a22 = h00 * h11 - h01 * h01;
    const double det = h00 * a00 + h01 * a01 + h02 * a02;
    d4* vec = MODE == P64_FWD_UPDATE ? pos : w;
    const d4 v0 = MODE == P64_FWD_UPDATE ? xs : ws;
    d4 vn = v0;
    if (fabs(det) > 1e-8) {
        const double inv = 1.0 / det;
        vn.x = v0.x + inv * (a00 * e0 + a01 * e1 + a02 * e2);
        vn.y = v0.y + inv * (a01 * e0 + a11 * e1 + a12 * e2);
        vn.z = v0.z + inv * (a02 * e0 + a12 * e1 + a22 * e2);
    }
    if constexpr (SAVE) save[s] = v0;
    vec[s] = vn;
}

template <int MODE, bool MAT, bool CONTACT, bool SAVE>
static hipError_t pass64_t(const DevMesh& m, const Dev64& d, int slot0, int count, d4* w, const d4* rhs, d4* out,
                           double* r2, d4* save, double mu, double lambda, double sdt, const ContactArgs& ct,
                           hipStream_t s) {
    const int per = 256 / LPV64;
    hipLaunchKernelGGL((k_pass64<MODE, MAT, CONTACT, SAVE>), dim3((count + per - 1) / per), dim3(256), 0, s, slot0, count,
                       d.pos, d.inert, m.mass, m.row, m.ptet, m.pa, m.pb, m.pc, m.pmat, w, rhs, out, r2, save, mu,
                       lambda, sdt, ct);
    return hipGetLastError();
}

template <int MODE>
static hipError_t pass64_m(const DevMesh& m, const Dev64& d, int slot0, int count, d4* w, const d4* rhs, d4* out,
                           double* r2, d4* save, double mu, double lambda, double sdt, const ContactArgs& ct,
                           hipStream_t s) {
    const bool mat = m.perTetMat, con = ct.on != 0, sv = save != nullptr;
#define VBD_P64(MT, CT, SV) \
    if (mat == MT && con == CT && sv == SV) \
        return pass64_t<MODE, MT, CT, SV>(m, d, slot0, count, w, rhs, out, r2, save, mu, lambda, sdt, ct, s);
    if constexpr (MODE == P64_FWD_UPDATE || MODE == P64_ADJ_SOLVE) {
        VBD_P64(false, false, true) VBD_P64(false, true, true) VBD_P64(true, false, true) VBD_P64(true, true, true)
    }
    VBD_P64(false, false, false) VBD_P64(false, true, false) VBD_P64(true, false, false) VBD_P64(true, true, false)
#undef VBD_P64
    return hipErrorInvalidValue;
}

static hipError_t pass64_t_diag(const DevMesh& m, const Dev64& d, int slot0, int count, d4* out, d4* save2, double mu,
                                double lambda, double sdt, const ContactArgs& ct, hipStream_t s) {
    const bool mat = m.perTetMat, con = ct.on != 0;
#define VBD_P64D(MT, CT) \
    if (mat == MT && con == CT) \
        return pass64_t<P64_DIAG_INV, MT, CT, false>(m, d, slot0, count, nullptr, nullptr, out, nullptr, save2, mu, \
                                                     lambda, sdt, ct, s);
    VBD_P64D(false, false) VBD_P64D(false, true) VBD_P64D(true, false) VBD_P64D(true, true)
#undef VBD_P64D
    return hipErrorInvalidValue;
}

hipError_t launch_pass64(int mode, const DevMesh& m, const Dev64& d, int slot0, int count, d4* w, const d4* rhs, d4* out,
                         double* r2, d4* save, double mu, double lambda, double sdt, const ContactArgs& ct,
                         hipStream_t s) {
    if (count <= 0) return hipSuccess;
    if (!m.pa || !d.pos) return hipErrorInvalidValue;  // full records and the FP64 state must exist
    switch (mode) {
        case P64_FWD_UPDATE: return pass64_m<P64_FWD_UPDATE>(m, d, slot0, count, w, rhs, out, r2, save, mu, lambda, sdt, ct, s);
        case P64_FWD_RESID: return pass64_m<P64_FWD_RESID>(m, d, slot0, count, w, rhs, out, r2, save, mu, lambda, sdt, ct, s);
        case P64_ADJ_SOLVE: return pass64_m<P64_ADJ_SOLVE>(m, d, slot0, count, w, rhs, out, r2, save, mu, lambda, sdt, ct, s);
        case P64_ADJ_RESID: return pass64_m<P64_ADJ_RESID>(m, d, slot0, count, w, rhs, out, r2, save, mu, lambda, sdt, ct, s);
        case P64_HVP: return pass64_m<P64_HVP>(m, d, slot0, count, w, rhs, out, r2, save, mu, lambda, sdt, ct, s);
        case P64_DIAG_INV:
            if (!out || !save) return hipErrorInvalidValue;
            return pass64_t_diag(m, d, slot0, count, out, save, mu, lambda, sdt, ct, s);
        default: return hipErrorInvalidValue;
    }
}

// forwardStep in double (VBD.cpp:103-112): x_prev = x; free: v += h g, y = x + h v; pinned: y = x
__global__ __launch_bounds__(256) void k_predict64(int n, int nFree, const d4* __restrict__ pos, d4* __restrict__ vel,
                                                   d4* __restrict__ xprev, d4* __restrict__ inert, double h, double gx,
                                                   double gy, double gz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const d4 x = pos[i];
    xprev[i] = x;
    d4 y = x;
    if (i < nFree) {
        d4 v = vel[i];
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

// updateVelocity in double (TetMesh.h:305-311): free: v = (x - x_prev) / h * damping
__global__ __launch_bounds__(256) void k_velocity64(int nFree, const d4* __restrict__ pos, const d4* __restrict__ xprev,
                                                    d4* __restrict__ vel, double h, double damp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nFree) return;
    const d4 x = pos[i], xp = xprev[i];
    d4 v = vel[i];
    v.x = (x.x - xp.x) / h * damp;
    v.y = (x.y - xp.y) / h * damp;
    v.z = (x.z - xp.z) / h * damp;
    vel[i] = v;
}

// x = omega (x - x_pp) + x_pp over the first n slots (VBD.cpp:290-300)
__global__ __launch_bounds__(256) void k_blend64(int n, d4* __restrict__ v, const d4* __restrict__ pp, double omega) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const d4 x = v[i], q = pp[i];
    d4 r = x;
    r.x = omega * (x.x - q.x) + q.x;
    r.y = omega * (x.y - q.y) + q.y;
    r.z = omega * (x.z - q.z) + q.z;
    v[i] = r;
}

__global__ __launch_bounds__(256) void k_copy_d4(int n, const d4* __restrict__ src, d4* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_f4_to_d4(int n, const f4* __restrict__ src, d4* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f4 a = src[i];
    dst[i] = d4{(double)a.x, (double)a.y, (double)a.z, 0.0};
}

__global__ __launch_bounds__(256) void k_d4_to_f4(int n, const d4* __restrict__ src, f4* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const d4 a = src[i];
    dst[i] = f4{(float)a.x, (float)a.y, (float)a.z, 0.f};
}

__global__ __launch_bounds__(256) void k_norm2_d4(int n, const d4* __restrict__ v, double* __restrict__ r2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const d4 a = v[i];
    r2[i] = a.x * a.x + a.y * a.y + a.z * a.z;
}

// caller xyz doubles (caller numbering) <-> slot double4; scatter_f: caller xyz floats -> slot double4
__global__ __launch_bounds__(256) void k_scatter_d(int n, const double* __restrict__ src3, const int* __restrict__ inv,
                                                   d4* __restrict__ dst) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    dst[inv[v]] = d4{src3[3 * (size_t)v], src3[3 * (size_t)v + 1], src3[3 * (size_t)v + 2], 0.0};
}

__global__ __launch_bounds__(256) void k_scatter_fd(int n, const float* __restrict__ src3, const int* __restrict__ inv,
                                                    d4* __restrict__ dst) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    dst[inv[v]] = d4{(double)src3[3 * (size_t)v], (double)src3[3 * (size_t)v + 1], (double)src3[3 * (size_t)v + 2], 0.0};
}

__global__ __launch_bounds__(256) void k_gather_d(int n, const d4* __restrict__ src, const int* __restrict__ inv,
                                                  double* __restrict__ dst3) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const d4 r = src[inv[v]];
    dst3[3 * (size_t)v] = r.x;
    dst3[3 * (size_t)v + 1] = r.y;
    dst3[3 * (size_t)v + 2] = r.z;
}

__global__ __launch_bounds__(256) void k_gather_df(int n, const d4* __restrict__ src, const int* __restrict__ inv,
                                                   float* __restrict__ dst3) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const d4 r = src[inv[v]];
    dst3[3 * (size_t)v] = (float)r.x;
    dst3[3 * (size_t)v + 1] = (float)r.y;
    dst3[3 * (size_t)v + 2] = (float)r.z;
}

// Per-tet -(d^2E / dx dtheta)^T w in double: dmu_t = -V (F - C) : dF(w), dlambda_t = -V (J-1) C : dF(w)
template <class T>
__global__ __launch_bounds__(256) void k_param_grad64(int nT, const i4* __restrict__ tets, const d4* __restrict__ pos,
                                                      const d4* __restrict__ w, const float* __restrict__ dminv,
                                                      const float* __restrict__ vol, T* __restrict__ dmu,
                                                      T* __restrict__ dlam) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const i4 tt = tets[t];
    const d4 x0 = pos[tt.x], x1 = pos[tt.y], x2 = pos[tt.z], x3 = pos[tt.w];
    const d4 w0 = w[tt.x], w1 = w[tt.y], w2 = w[tt.z], w3 = w[tt.w];
    const float* D = dminv + 9 * (size_t)t;
    const double ds[3][3] = {{x1.x - x0.x, x2.x - x0.x, x3.x - x0.x},
                             {x1.y - x0.y, x2.y - x0.y, x3.y - x0.y},
                             {x1.z - x0.z, x2.z - x0.z, x3.z - x0.z}};
    const double dw[3][3] = {{w1.x - w0.x, w2.x - w0.x, w3.x - w0.x},
                             {w1.y - w0.y, w2.y - w0.y, w3.y - w0.y},
                             {w1.z - w0.z, w2.z - w0.z, w3.z - w0.z}};
    double F[3][3], G[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            F[r][c] = ds[r][0] * D[c] + ds[r][1] * D[3 + c] + ds[r][2] * D[6 + c];
            G[r][c] = dw[r][0] * D[c] + dw[r][1] * D[3 + c] + dw[r][2] * D[6 + c];
        }
    double C[3][3];
    C[0][0] = F[1][1] * F[2][2] - F[1][2] * F[2][1];
    C[0][1] = F[1][2] * F[2][0] - F[1][0] * F[2][2];
    C[0][2] = F[1][0] * F[2][1] - F[1][1] * F[2][0];
    C[1][0] = F[2][1] * F[0][2] - F[2][2] * F[0][1];
    C[1][1] = F[2][2] * F[0][0] - F[2][0] * F[0][2];
    C[1][2] = F[2][0] * F[0][1] - F[2][1] * F[0][0];
    C[2][0] = F[0][1] * F[1][2] - F[0][2] * F[1][1];
    C[2][1] = F[0][2] * F[1][0] - F[0][0] * F[1][2];
    C[2][2] = F[0][0] * F[1][1] - F[0][1] * F[1][0];
    const double J = F[0][0] * C[0][0] + F[0][1] * C[0][1] + F[0][2] * C[0][2];
    double fc = 0.0, cg = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            fc += (F[r][c] - C[r][c]) * G[r][c];
            cg += C[r][c] * G[r][c];
        }
    dmu[t] = (T)(-(double)vol[t] * fc);
    dlam[t] = (T)(-(double)vol[t] * (J - 1.0) * cg);
}

// ---------------------------------------------------------------------------------------
// Conjugate gradients on A = M/h^2 + K (+ contact) with the block-Jacobi preconditioner D^-1
// (the inverses of the 3x3 diagonal blocks, P64_DIAG_INV), in the single-reduction form of
// Chronopoulos and Gear: with s = A p carried by the recurrence s = w + beta s (w = A z), the
// two dot products of an iteration, gamma = r.z and delta = w.z, come from one reduction, so an
// iteration is three launches: the update below, the A z pass (with w.z per slot) and one
// launch summing gamma, delta and r.r.  Iteration k:
//   beta_k = gamma_k / gamma_{k-1} (0 for k = 0), alpha_k = gamma_k / (delta_k - beta_k gamma_k / alpha_{k-1})
//   p = z + beta p, s = w + beta s, x += alpha p, r -= alpha s, z = D^-1 r  (then w = A z)
// The scalars stay on the device: cg_scal[CG_G + k&1] = gamma, [CG_D] = delta, [CG_A + k&1] =
// alpha, [CG_RR] = r.r, [CG_BAD] set when gamma or the alpha denominator is not positive.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ d4 dinv_apply(const d4 a, const d4 b, const d4 r) {
    // [[a.x a.y a.z] [a.y a.w b.x] [a.z b.x b.y]] r
    return d4{a.x * r.x + a.y * r.y + a.z * r.z, a.y * r.x + a.w * r.y + b.x * r.z,
              a.z * r.x + b.x * r.y + b.y * r.z, 0.0};
}

__device__ __forceinline__ double dot3(const d4 a, const d4 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// z = D^-1 r, p = s = 0, part_rz = r.z, part_rr = r.r (the start: r = rhs - A x)
__global__ __launch_bounds__(256) void k_cg_init64(int n, const d4* __restrict__ r, d4* __restrict__ z,
                                                   d4* __restrict__ p, d4* __restrict__ sv, const d4* __restrict__ da,
                                                   const d4* __restrict__ db, double* __restrict__ partRz,
                                                   double* __restrict__ partRr, double* __restrict__ scal) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) scal[CG_BAD] = 0.0;
    if (i >= n) return;
    const d4 ri = r[i];
    const d4 zi = dinv_apply(da[i], db[i], ri);
    z[i] = zi;
    p[i] = d4{0.0, 0.0, 0.0, 0.0};
    sv[i] = d4{0.0, 0.0, 0.0, 0.0};
    partRz[i] = dot3(ri, zi);
    partRr[i] = dot3(ri, ri);
}

// iteration k (see above; par = k & 1, first = (k == 0)): directions, x and r, the new z, and
// r.z, r.r per slot
__global__ __launch_bounds__(256) void k_cg_update64(int n, int par, int first, d4* __restrict__ x, d4* __restrict__ r,
                                                     d4* __restrict__ z, d4* __restrict__ p, d4* __restrict__ sv,
                                                     const d4* __restrict__ w, const d4* __restrict__ da,
                                                     const d4* __restrict__ db, double* __restrict__ scal,
                                                     double* __restrict__ partRz, double* __restrict__ partRr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double gamma = scal[CG_G + par], delta = scal[CG_D];
    double beta = 0.0, denom = delta;
    if (!first) {
        beta = gamma / scal[CG_G + (par ^ 1)];
        denom = delta - beta * gamma / scal[CG_A + (par ^ 1)];
    }
    const double alpha = gamma / denom;
    if (i == 0) {
        scal[CG_A + par] = alpha;
        if (!(gamma > 0.0 && denom > 0.0)) scal[CG_BAD] = 1.0;
    }
    if (i >= n) return;
    const d4 zi = z[i], wi = w[i];
    d4 pi = p[i], si = sv[i], xi = x[i], ri = r[i];
    pi.x = zi.x + beta * pi.x; pi.y = zi.y + beta * pi.y; pi.z = zi.z + beta * pi.z;
    si.x = wi.x + beta * si.x; si.y = wi.y + beta * si.y; si.z = wi.z + beta * si.z;
    xi.x += alpha * pi.x; xi.y += alpha * pi.y; xi.z += alpha * pi.z;
    ri.x -= alpha * si.x; ri.y -= alpha * si.y; ri.z -= alpha * si.z;
    p[i] = pi;
    sv[i] = si;
    x[i] = xi;
    r[i] = ri;
    const d4 zn = dinv_apply(da[i], db[i], ri);
    z[i] = zn;
    partRz[i] = dot3(ri, zn);
    partRr[i] = dot3(ri, ri);
}

// Fixed-order sums of three arrays in one launch, block b summing array b exactly as
// k_sum_fixed does (the same order, so the same bits)
__global__ __launch_bounds__(1024) void k_sum3_fixed(int n, const double* __restrict__ v0, double* __restrict__ o0,
                                                     const double* __restrict__ v1, double* __restrict__ o1,
                                                     const double* __restrict__ v2, double* __restrict__ o2) {
    __shared__ double sh[1024];
    const double* v = blockIdx.x == 0 ? v0 : (blockIdx.x == 1 ? v1 : v2);
    double* out = blockIdx.x == 0 ? o0 : (blockIdx.x == 1 ? o1 : o2);
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

// dst = a src (the Newton right-hand side -grad G)
__global__ __launch_bounds__(256) void k_scale_d4(int n, const d4* __restrict__ src, d4* __restrict__ dst, double a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const d4 v = src[i];
    dst[i] = d4{a * v.x, a * v.y, a * v.z, 0.0};
}

// pos = base + alpha dx (a Newton step of length alpha from base)
__global__ __launch_bounds__(256) void k_newton_move64(int n, d4* __restrict__ pos, const d4* __restrict__ base,
                                                       const d4* __restrict__ dx, double alpha) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const d4 b = base[i], d = dx[i];
    d4 x = pos[i];
    x.x = b.x + alpha * d.x;
    x.y = b.y + alpha * d.y;
    x.z = b.z + alpha * d.z;
    pos[i] = x;
}

static inline dim3 grid256(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_scale_d4(int n, const d4* src, d4* dst, double a, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale_d4, grid256(n), dim3(256), 0, s, n, src, dst, a);
    return hipGetLastError();
}

hipError_t launch_newton_move64(int n, d4* pos, const d4* base, const d4* dx, double alpha, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_newton_move64, grid256(n), dim3(256), 0, s, n, pos, base, dx, alpha);
    return hipGetLastError();
}

hipError_t launch_cg_init64(int n, const Dev64& d, double* partRz, double* partRr, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_init64, grid256(n > 0 ? n : 1), dim3(256), 0, s, n, d.cg_r, d.cg_z, d.cg_p, d.cg_s, d.cg_da,
                       d.cg_db, partRz, partRr, d.cg_scal);
    return hipGetLastError();
}

hipError_t launch_cg_update64(int n, int k, const Dev64& d, d4* x, const d4* w, double* partRz, double* partRr,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_cg_update64, grid256(n > 0 ? n : 1), dim3(256), 0, s, n, k & 1, k == 0 ? 1 : 0, x, d.cg_r,
                       d.cg_z, d.cg_p, d.cg_s,
                       w, d.cg_da, d.cg_db, d.cg_scal, partRz, partRr);
    return hipGetLastError();
}

hipError_t launch_sum3_fixed(int n, const double* v0, double* o0, const double* v1, double* o1, const double* v2,
                             double* o2, hipStream_t s) {
    if (n > SUM_TWO_STAGE) return hipErrorInvalidValue;  // the callers sum larger arrays in two stages
    hipLaunchKernelGGL(k_sum3_fixed, dim3(3), dim3(1024), 0, s, n, v0, o0, v1, o1, v2, o2);
    return hipGetLastError();
}

hipError_t launch_predict64(const DevMesh& m, const Dev64& d, double h, const float g[3], hipStream_t s) {
    if (m.nV == 0) return hipSuccess;
    hipLaunchKernelGGL(k_predict64, grid256(m.nV), dim3(256), 0, s, m.nV, m.nFree, d.pos, d.vel, d.xprev, d.inert, h,
                       (double)g[0], (double)g[1], (double)g[2]);
    return hipGetLastError();
}

hipError_t launch_velocity64(const DevMesh& m, const Dev64& d, double h, double damping, hipStream_t s) {
    if (m.nFree == 0) return hipSuccess;
    hipLaunchKernelGGL(k_velocity64, grid256(m.nFree), dim3(256), 0, s, m.nFree, d.pos, d.xprev, d.vel, h, damping);
    return hipGetLastError();
}

hipError_t launch_blend64(int n, d4* v, const d4* pp, double omega, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_blend64, grid256(n), dim3(256), 0, s, n, v, pp, omega);
    return hipGetLastError();
}

hipError_t launch_copy_d4(int n, const d4* src, d4* dst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_d4, grid256(n), dim3(256), 0, s, n, src, dst);
    return hipGetLastError();
}

hipError_t launch_f4_to_d4(int n, const f4* src, d4* dst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_f4_to_d4, grid256(n), dim3(256), 0, s, n, src, dst);
    return hipGetLastError();
}

hipError_t launch_d4_to_f4(int n, const d4* src, f4* dst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_d4_to_f4, grid256(n), dim3(256), 0, s, n, src, dst);
    return hipGetLastError();
}

hipError_t launch_norm2_d4(int n, const d4* v, double* r2, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_norm2_d4, grid256(n), dim3(256), 0, s, n, v, r2);
    return hipGetLastError();
}

hipError_t launch_scatter_d(int n, const double* src3, const int* inv, d4* dst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_d, grid256(n), dim3(256), 0, s, n, src3, inv, dst);
    return hipGetLastError();
}

hipError_t launch_scatter_fd(int n, const float* src3, const int* inv, d4* dst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_fd, grid256(n), dim3(256), 0, s, n, src3, inv, dst);
    return hipGetLastError();
}

hipError_t launch_gather_d(int n, const d4* src, const int* inv, double* dst3, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_d, grid256(n), dim3(256), 0, s, n, src, inv, dst3);
    return hipGetLastError();
}

hipError_t launch_gather_df(int n, const d4* src, const int* inv, float* dst3, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_df, grid256(n), dim3(256), 0, s, n, src, inv, dst3);
    return hipGetLastError();
}

hipError_t launch_param_grad64(const DevMesh& m, const Dev64& d, const d4* w, float* dmu, float* dlam, hipStream_t s) {
    if (m.nT == 0) return hipSuccess;
    hipLaunchKernelGGL(k_param_grad64<float>, grid256(m.nT), dim3(256), 0, s, m.nT, m.tets, d.pos, w, m.dminv, m.vol,
                       dmu, dlam);
    return hipGetLastError();
}

hipError_t launch_param_grad64(const DevMesh& m, const Dev64& d, const d4* w, double* dmu, double* dlam,
                               hipStream_t s) {
    if (m.nT == 0) return hipSuccess;
    hipLaunchKernelGGL(k_param_grad64<double>, grid256(m.nT), dim3(256), 0, s, m.nT, m.tets, d.pos, w, m.dminv, m.vol,
                       dmu, dlam);
    return hipGetLastError();
}

hipError_t alloc_cg64(Dev64& d, int nV) {
    const size_t n = (size_t)(nV > 0 ? nV : 1);
    d4** arrs[] = {&d.cg_r, &d.cg_z, &d.cg_p, &d.cg_s, &d.cg_da, &d.cg_db};
    for (d4** a : arrs) {
        if (*a) continue;
        hipError_t e = hipMalloc((void**)a, sizeof(d4) * n);
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipSuccess;
    if (!d.cg_part && (e = hipMalloc((void**)&d.cg_part, sizeof(double) * 2 * (n + SUM_STAGE))) != hipSuccess)
        return e;
    if (!d.cg_scal && (e = hipMalloc((void**)&d.cg_scal, sizeof(double) * CG_SCALARS)) != hipSuccess) return e;
    return e;
}

hipError_t alloc_dev64(Dev64& d, int nV) {
    const size_t n = (size_t)(nV > 0 ? nV : 1);
    d4** arrs[] = {&d.pos, &d.inert, &d.vel, &d.xprev, &d.save[0], &d.save[1],
                   &d.w, &d.rhs, &d.out, &d.wsave[0], &d.wsave[1]};
    for (d4** a : arrs) {
        if (*a) continue;
        hipError_t e = hipMalloc((void**)a, sizeof(d4) * n);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

void free_dev64(Dev64& d) {
    d4* arrs[] = {d.pos, d.inert, d.vel, d.xprev, d.save[0], d.save[1], d.w, d.rhs, d.out, d.wsave[0], d.wsave[1],
                  d.cg_r, d.cg_z, d.cg_p, d.cg_s, d.cg_da, d.cg_db};
    for (d4* a : arrs)
        if (a) (void)hipFree(a);
    if (d.cg_part) (void)hipFree(d.cg_part);
    if (d.cg_scal) (void)hipFree(d.cg_scal);
    d = Dev64();
}

}  // namespace vbd
