// fs_kernels.hip -- gfx950 kernels of the shell finite-element hot path.
//
// Element stiffness (T3FF / T3FFComp / Q4RS / Q4RSComp)
// ----------------------------------------------------
// The reference forms K_e in the element frame from B matrices
// (add_btdb_ut_only!, FEMMShellT3FFModule.jl:680-706; FEMMShellQ4RSModule.jl:
// 929-936), then applies two dense 18x18 (24x24) congruences
// (TransformerQtEQ, :711-730).  Here the same matrix is formed as
//
//     K_e = sum_q C_q^T C_q  (+ drilling),   C_q = L_q * B_q * T_ae * T_ga
//
// where L_q is the upper Cholesky factor of the (scaled) material matrix at
// quadrature point q (Dps, Dt; or [A B; B D], H rotated by the layup angle),
// B_q stacks the membrane, bending and transverse-shear strain-displacement
// rows, and T_ae * T_ga is the nodal-to-element times global-to-nodal
// transformation.  C_q has 8 rows (12 for the T3 AVERAGE_K formulation) and
// 6*nen columns; with exact arithmetic this equals the reference's
// T_ga' (T_ae' K_loc T_ae + drill) T_ga.  The T3 drilling stiffness (added in
// the nodal basis, :714-727) becomes a rank-one term kavg * g g' on the
// rotation block, with g the nodal drilling axis in the global basis; the Q4
// drilling (FEMMShellQ4RSModule.jl:807-859) is applied to the assembled
// global K_e exactly as the reference does.
//
// One workgroup processes NE elements in phases separated by barriers:
//   G1  thread per (element, quadrature point): geometry, scaled material
//       factors, shear B-matrix entries                        -> LDS
//   G2  thread per (element, qp, node): nodal triad (Rodrigues), C columns
//       of the node in the element frame, coupling vector d   -> LDS
//   G3  thread per (element, qp, node): global-basis C columns -> LDS
//   C   thread per (element, upper node pair): 6x6 block = C_a^T C_b summed
//       over all rows of all quadrature points, plus drilling
//   W   blocks staged in LDS and stored contiguously (Kb[e][blk][36]).
//
// Deterministic scatter into CSR
// ------------------------------
// scatter_pair_kernel sums, for each upper node pair (I, J), the
// contributing element blocks in ascending element order (the order in which
// sparse(I, J, V) combines duplicates) and writes the (I, J) and (J, I)
// 6x6 blocks to their CSR positions.  No atomics: every CSR value is written
// exactly once and the result is bitwise reproducible.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>

#include "fs_device.h"

namespace fs {

// ---------------------------------------------------------------------------
// Small dense helpers (FP64)
// ---------------------------------------------------------------------------

// _e_g! (FEMMShellT3FFModule.jl:283-305): element basis from the tangents.
__device__ __forceinline__ void e_g(double EG[3][3], const double J[3][2]) {
    EG[0][0] = J[0][0];
    EG[1][0] = J[1][0];
    EG[2][0] = J[2][0];
    double n = 1.0 / sqrt(EG[0][0] * EG[0][0] + EG[1][0] * EG[1][0] + EG[2][0] * EG[2][0]);
    EG[0][0] *= n;
    EG[1][0] *= n;
    EG[2][0] *= n;
    EG[0][2] = -EG[2][0] * J[1][1] + EG[1][0] * J[2][1];
    EG[1][2] = EG[2][0] * J[0][1] - EG[0][0] * J[2][1];
    EG[2][2] = -EG[1][0] * J[0][1] + EG[0][0] * J[1][1];
    n = 1.0 / sqrt(EG[0][2] * EG[0][2] + EG[1][2] * EG[1][2] + EG[2][2] * EG[2][2]);
    EG[0][2] *= n;
    EG[1][2] *= n;
    EG[2][2] *= n;
    EG[0][1] = -EG[2][2] * EG[1][0] + EG[1][2] * EG[2][0];
    EG[1][1] = EG[2][2] * EG[0][0] - EG[0][2] * EG[2][0];
    EG[2][1] = -EG[1][2] * EG[0][0] + EG[0][2] * EG[1][0];
}

// Layup2ElementAngle (TransformerModule.jl:92-105): cos/sin of the angle
// between the element frame and the layup reference direction.  The
// argument of the square root is clamped at zero (the reference would raise
// a DomainError for a rounding excursion of m above one).
__device__ __forceinline__ void layup_angle(const double EG[3][3], const double *L, double &m,
                                            double &n) {
    double M[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            M[i][j] = EG[0][i] * L[0 * 3 + j] + EG[1][i] * L[1 * 3 + j] + EG[2][i] * L[2 * 3 + j];
    double c0 = sqrt(M[0][0] * M[0][0] + M[1][0] * M[1][0]);
    double c1 = sqrt(M[0][1] * M[0][1] + M[1][1] * M[1][1]);
    double m11 = M[0][0] / c0, m21 = M[1][0] / c0, m12 = M[0][1] / c1, m22 = M[1][1] / c1;
    double mm = (m11 + m22) / 2;
    double nn = (m12 - m21) / 2;
    double sn = (nn >= 0.0) ? 1.0 : -1.0;
    double arg = 1 - mm * mm;
    m = mm;
    n = sn * sqrt(arg > 0.0 ? arg : 0.0);
}

// Effective material factors of a layered shell at one point:
//   Lmb = s_mb * Labd * diag(Tb, Tb),  Ls = s_s * Lh * Tt
// with Tb = plane_stress_Tbar_matrix!(m, n) (CompositeLayupModule.jl:364-378)
// and Tt = transverse_shear_T_matrix!(m, n) (:463-471), so that
// Lmb' Lmb = s_mb^2 [Tb' A Tb, Tb' B Tb; Tb' B Tb, Tb' D Tb] and
// Ls' Ls = s_s^2 Tt' H Tt -- the sA, sB, sD, sH of FEMMShellT3FFCompModule.jl:614-626.
__device__ __forceinline__ void composite_factors(const double *Labd, const double *Lh, double m, double n,
                                                  double s_mb, double s_s, double Lmb[36],
                                                  double Ls[4]) {
    const double Tb[3][3] = {{m * m, n * n, -(m * n)}, {n * n, m * m, (m * n)},
                             {2 * (m * n), -2 * (m * n), m * m - n * n}};
    for (int i = 0; i < 6; ++i)
        for (int blk = 0; blk < 2; ++blk)
            for (int j = 0; j < 3; ++j) {
                double acc = 0.0;
                for (int k = 0; k < 3; ++k) acc += Labd[i * 6 + blk * 3 + k] * Tb[k][j];
                Lmb[i * 6 + blk * 3 + j] = s_mb * acc;
            }
    const double Tt[2][2] = {{m, -n}, {n, m}};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) Ls[i * 2 + j] = s_s * (Lh[i * 2 + 0] * Tt[0][j] + Lh[i * 2 + 1] * Tt[1][j]);
}

// Homogeneous factors: Lmb = diag(s_m Lps, s_b Lps), Ls = s_s Lt.
__device__ __forceinline__ void homog_factors(const double *Lps, const double *Lt, double s_m, double s_b,
                                              double s_s, double Lmb[36], double Ls[4]) {
    for (int i = 0; i < 36; ++i) Lmb[i] = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Lmb[i * 6 + j] = s_m * Lps[i * 3 + j];
            Lmb[(3 + i) * 6 + 3 + j] = s_b * Lps[i * 3 + j];
        }
    for (int i = 0; i < 4; ++i) Ls[i] = s_s * Lt[i];
}

// XCD-aware block order.  Workgroups are dispatched round-robin over the 8
// XCDs (block b runs on XCD b mod 8) and every XCD has its own L2.  Renumber
// so that each XCD processes one contiguous range of the logical blocks:
// neighbouring rows / nodes then share an L2, and the gathers of a sparse
// row (u, coordinates, element factors) hit there instead of HBM.  A
// bijection for any block count.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nblocks) {
    const int64_t q = nblocks / 8, r = nblocks % 8;
    const int64_t x = b % 8, k = b / 8;
    return x * q + (x < r ? x : r) + k;
}

// _NodalTriadsE (FEMMShellT3FFModule.jl:355-388): rotation taking the
// element normal e3 to the nodal normal nk_e (element frame); the rotation
// vector is r = e3 x nk_e and its length |r| = sin(theta) is used as the
// angle (rotmat3!), identity below 1e-12.
__device__ __forceinline__ void nodal_triad(const double nk_e[3], double A[3][3]) {
    const double r0 = -nk_e[1], r1 = nk_e[0];
    const double na = sqrt(r0 * r0 + r1 * r1);
    if (na > 1.0e-12) {
        const double ina = 1.0 / na;
        const double u0 = r0 * ina, u1 = r1 * ina;
        double sa, ca;
        sincos(na, &sa, &ca);
        const double f = sa * ina;
        const double uu[3] = {u0, u1, 0.0};
        const double Sk[3][3] = {{0.0, 0.0, r1}, {0.0, 0.0, -r0}, {-r1, r0, 0.0}};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double aa = uu[i] * uu[j];
                A[i][j] = ca * (((i == j) ? 1.0 : 0.0) - aa) + f * Sk[i][j] + aa;
            }
    } else {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) A[i][j] = (i == j) ? 1.0 : 0.0;
    }
}

// Material factor rows at one quadrature point, produced on demand (the
// rows of Lmb and Ls of homog_factors / composite_factors, same arithmetic).
struct MatRows {
    bool composite;
    const double *L;  // chol(Dps) (homogeneous) or chol([A B; B D]) (layered)
    double s_m, s_b;  // membrane / bending scales (layered: s_m = s_mb)
    double m, n;      // layup angle cosine / sine (layered)
    double Ls[4];
};

__device__ __forceinline__ MatRows mat_rows(const ElemParams &p, bool composite, int grp, double sc0, double sc1,
                                            double sc2, double sc3, double sc4) {
    MatRows r;
    r.composite = composite;
    if (!r.composite) {
        const double *hrec = p.g_homog + 16 * grp;
        r.L = hrec;
        r.s_m = sc0;
        r.s_b = sc1;
        r.m = r.n = 0.0;
        for (int i = 0; i < 4; ++i) r.Ls[i] = sc2 * hrec[9 + i];
    } else {
        r.L = p.g_Labd + 36 * grp;
        r.s_m = sc0;
        r.s_b = 0.0;
        r.m = sc3;
        r.n = sc4;
        const double *Lh = p.g_Lh + 4 * grp;
        const double Tt[2][2] = {{sc3, -sc4}, {sc4, sc3}};
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) r.Ls[i * 2 + j] = sc2 * (Lh[i * 2 + 0] * Tt[0][j] + Lh[i * 2 + 1] * Tt[1][j]);
    }
    return r;
}

__device__ __forceinline__ void lmb_row(const MatRows &mr, int s, double row[6]) {
    if (!mr.composite) {
        for (int j = 0; j < 6; ++j) row[j] = 0.0;
        if (s < 3) {
            for (int j = 0; j < 3; ++j) row[j] = mr.s_m * mr.L[s * 3 + j];
        } else {
            for (int j = 0; j < 3; ++j) row[3 + j] = mr.s_b * mr.L[(s - 3) * 3 + j];
        }
    } else {
        const double m = mr.m, n = mr.n;
        const double Tb[3][3] = {{m * m, n * n, -(m * n)}, {n * n, m * m, (m * n)},
                                 {2 * (m * n), -2 * (m * n), m * m - n * n}};
        for (int blk = 0; blk < 2; ++blk)
            for (int j = 0; j < 3; ++j) {
                double acc = 0.0;
                for (int k = 0; k < 3; ++k) acc += mr.L[s * 6 + blk * 3 + k] * Tb[k][j];
                row[blk * 3 + j] = mr.s_m * acc;
            }
    }
}

// Element-frame columns (u, v, w, theta_x, theta_y) of one node for row s of
// C: membrane/bending rows from a row of Lmb and the shape-function
// gradients, shear rows from Ls and the node's shear B entries
// bs[comp * NSH + row] (comp 0 = w, 1 = theta_x, 2 = theta_y).
__device__ __forceinline__ void c_row_mb(const double lrow[6], double gN0, double gN1, double &cu, double &cv,
                                         double &cw, double &cx, double &cy) {
    cu = lrow[0] * gN0 + lrow[2] * gN1;
    cv = lrow[1] * gN1 + lrow[2] * gN0;
    cw = 0.0;
    cx = -lrow[4] * gN1 - lrow[5] * gN0;
    cy = lrow[3] * gN0 + lrow[5] * gN1;
}

template <int NSH>
__device__ __forceinline__ void c_row_sh(const double Ls[4], int s, const double *bs, double &cu, double &cv,
                                         double &cw, double &cx, double &cy) {
    const int o = (s - 6) / 2, rr = (s - 6) % 2;
    // select the factor row without indexing (keeps Ls in registers)
    const double l0 = rr ? Ls[2] : Ls[0], l1 = rr ? Ls[3] : Ls[1];
    cu = 0.0;
    cv = 0.0;
    cw = l0 * bs[0 * NSH + 2 * o] + l1 * bs[0 * NSH + 2 * o + 1];
    cx = l0 * bs[1 * NSH + 2 * o] + l1 * bs[1 * NSH + 2 * o + 1];
    cy = l0 * bs[2 * NSH + 2 * o] + l1 * bs[2 * NSH + 2 * o + 1];
}

__device__ __forceinline__ void c_row_q4(const MatRows &mr, int s, double gN0, double gN1, const double *bs,
                                         double &cu, double &cv, double &cw, double &cx, double &cy) {
    if (s < 6) {
        double lrow[6];
        lmb_row(mr, s, lrow);
        c_row_mb(lrow, gN0, gN1, cu, cv, cw, cx, cy);
    } else {
        c_row_sh<2>(mr.Ls, s, bs, cu, cv, cw, cx, cy);
    }
}

// Row of C handled by iteration i of lane half h when the eight Q4 rows of
// a node are split over two lanes: iterations 0..2 take the membrane rows
// (h = 0) or the bending rows (h = 1) built from row i of the material
// factor, iteration 3 the shear row 6 + h.  The two halves follow the same
// control flow; their values equal those of c_row_q4 for the same row.
__device__ __forceinline__ int q4_split_row(int i, int h) { return i < 3 ? 3 * h + i : 6 + h; }

__device__ __forceinline__ void c_row_q4_split(const MatRows &mr, int i, int h, double gN0, double gN1,
                                               const double *bs, double &cu, double &cv, double &cw,
                                               double &cx, double &cy) {
    if (i < 3) {
        double lrow[6];
        if (!mr.composite) {
            const double sc = h ? mr.s_b : mr.s_m;
            for (int j = 0; j < 3; ++j) {
                const double v = sc * mr.L[i * 3 + j];
                lrow[j] = h ? 0.0 : v;
                lrow[3 + j] = h ? v : 0.0;
            }
        } else {
            lmb_row(mr, 3 * h + i, lrow);
        }
        c_row_mb(lrow, gN0, gN1, cu, cv, cw, cx, cy);
    } else {
        c_row_sh<2>(mr.Ls, 6 + h, bs, cu, cv, cw, cx, cy);
    }
}

// ---------------------------------------------------------------------------
// T3 element kernel
// ---------------------------------------------------------------------------

template <int NEN, int NQP, int NR>
struct ElemTraits {
    static constexpr int K = 6 * NEN;                  // element dofs
    static constexpr int NBLK = NEN * (NEN + 1) / 2;   // upper node-pair blocks
    static constexpr int NS = NQP * NR;                // rows of C per element
    static constexpr int CSTR = NS * K + 2;            // LDS stride of C (padded)
    static constexpr int NSH = NR - 6;                 // raw shear rows (2 or 6)
    // geometry record per (element, qp)
    static constexpr int G_EG = 0;                     // 9
    static constexpr int G_GN = 9;                     // 2 * NEN  (gradN_e)
    static constexpr int G_LMB = G_GN + 2 * NEN;       // 36
    static constexpr int G_LS = G_LMB + 36;            // 4
    static constexpr int G_BS = G_LS + 4;              // NEN * 3 * NSH: bs[node][comp w,tx,ty][row]
    static constexpr int GEO = G_BS + NEN * 3 * NSH;
    // node record per (element, qp, node): A rows 0..1 and coupling vector d
    static constexpr int NODE = 6 + NR;
    // per element: drilling vectors and the parts of the drilling average
    static constexpr int M_G = 0;                      // NEN*3
    static constexpr int M_KP = NEN * 3;               // NEN*2
    static constexpr int MISC = M_KP + NEN * 2;
    static constexpr int KSTAGE = NBLK * 36;           // staged output per element
};

template <int NEN, int NQP, int NR, int NE>
struct ElemSmem {
    using T = ElemTraits<NEN, NQP, NR>;
    // the output stage (NE x KSTAGE, after the last read of C and of the
    // geometry) spans the C and G regions, so C is sized for C alone where
    // the two together hold the stage
    static constexpr int SZ_G = NE * NQP * T::GEO;
    static constexpr int OFF_C = 0;
    static constexpr int SZ_C = NE * T::CSTR + SZ_G >= NE * T::KSTAGE ? NE * T::CSTR : NE * T::KSTAGE - SZ_G;
    static constexpr int OFF_G = SZ_C;
    static constexpr int OFF_N = OFF_G + SZ_G;
    static constexpr int SZ_N = NE * NQP * NEN * T::NODE;
    static constexpr int OFF_M = OFF_N + SZ_N;
    static constexpr int SZ_M = NE * T::MISC;
    static constexpr int TOTAL = OFF_M + SZ_M;  // doubles
};

// Q4 basis-function derivatives at (xi, eta) (FinEtools FESetQ4 bfundpar).
__device__ __forceinline__ void q4_gradn(double xi, double eta, double gN[4][2]) {
    gN[0][0] = -(1. - eta) * 0.25;
    gN[0][1] = -(1. - xi) * 0.25;
    gN[1][0] = (1. - eta) * 0.25;
    gN[1][1] = -(1. + xi) * 0.25;
    gN[2][0] = (1. + eta) * 0.25;
    gN[2][1] = (1. + xi) * 0.25;
    gN[3][0] = -(1. + eta) * 0.25;
    gN[3][1] = (1. - xi) * 0.25;
}

// GaussRule(2, 2) abscissa as FinEtools tabulates it.
#define FS_GAUSS2 0.577350269189626

// MITC transverse shear strain-displacement entries at (r, s) before the
// transformation (_Bsmat, FEMMShellQ4RSModule.jl:753-804): bs[node][comp
// w, tx, ty][row], from the element coordinates ec relative to the centroid.
__device__ __forceinline__ void q4_mitc_bs(const double ec[4][2], double r, double s, double *bs) {
    const double X1 = ec[0][0], Y1 = ec[0][1], X2 = ec[1][0], Y2 = ec[1][1];
    const double X3 = ec[2][0], Y3 = ec[2][1], X4 = ec[3][0], Y4 = ec[3][1];
    const double J11 = (X1 * (s - 1) / 4 - X2 * (s - 1) / 4 + X3 * (s + 1) / 4 - X4 * (s + 1) / 4);
    const double J21 = (Y1 * (s - 1) / 4 - Y2 * (s - 1) / 4 + Y3 * (s + 1) / 4 - Y4 * (s + 1) / 4);
    const double J12 = (X1 * (r - 1) / 4 - X2 * (r + 1) / 4 + X3 * (r + 1) / 4 - X4 * (r - 1) / 4);
    const double J22 = (Y1 * (r - 1) / 4 - Y2 * (r + 1) / 4 + Y3 * (r + 1) / 4 - Y4 * (r - 1) / 4);
    const double A = sqrt(J11 * J11 + J21 * J21);
    const double B = sqrt(J12 * J12 + J22 * J22);
    const double iA = 1.0 / A, iB = 1.0 / B;
    const double ca = J11 * iA, sa = J21 * iA, cb = J12 * iB, sb = J22 * iB;
    const double detJ = J11 * J22 - J12 * J21;
    const double Ax = X1 - X2 - X3 + X4, Ay = Y1 - Y2 - Y3 + Y4;
    const double Bx = X1 - X2 + X3 - X4, By = Y1 - Y2 + Y3 - Y4;
    const double Cx = X1 + X2 - X3 - X4, Cy = Y1 + Y2 - Y3 - Y4;
    const double P = sqrt((Ax + Bx * s) * (Ax + Bx * s) + (Ay + By * s) * (Ay + By * s));
    const double Q = sqrt((Bx * r + Cx) * (Bx * r + Cx) + (By * r + Cy) * (By * r + Cy));
    const double i16 = 1.0 / (16 * detJ), i32 = 0.5 * i16;  // exact: power-of-two scale
    bs[0 * 6 + 0 * 2 + 0] = (sa * (r + 1) * P - sb * (s + 1) * Q) * i16;
    bs[0 * 6 + 1 * 2 + 0] = (-sa * (Y1 - Y4) * (r + 1) * P + sb * (Y1 - Y2) * (s + 1) * Q) * i32;
    bs[0 * 6 + 2 * 2 + 0] = (sa * (X1 - X4) * (r + 1) * P - sb * (X1 - X2) * (s + 1) * Q) * i32;
    bs[0 * 6 + 0 * 2 + 1] = (-ca * (r + 1) * P + cb * (s + 1) * Q) * i16;
    bs[0 * 6 + 1 * 2 + 1] = (ca * (Y1 - Y4) * (r + 1) * P - cb * (Y1 - Y2) * (s + 1) * Q) * i32;
    bs[0 * 6 + 2 * 2 + 1] = (-ca * (X1 - X4) * (r + 1) * P + cb * (X1 - X2) * (s + 1) * Q) * i32;
    bs[1 * 6 + 0 * 2 + 0] = (-sa * (r - 1) * P + sb * (s + 1) * Q) * i16;
    bs[1 * 6 + 1 * 2 + 0] = (sa * (Y2 - Y3) * (r - 1) * P + sb * (Y1 - Y2) * (s + 1) * Q) * i32;
    bs[1 * 6 + 2 * 2 + 0] = (-sa * (X2 - X3) * (r - 1) * P - sb * (X1 - X2) * (s + 1) * Q) * i32;
    bs[1 * 6 + 0 * 2 + 1] = (ca * (r - 1) * P - cb * (s + 1) * Q) * i16;
    bs[1 * 6 + 1 * 2 + 1] = (-ca * (Y2 - Y3) * (r - 1) * P - cb * (Y1 - Y2) * (s + 1) * Q) * i32;
    bs[1 * 6 + 2 * 2 + 1] = (ca * (X2 - X3) * (r - 1) * P + cb * (X1 - X2) * (s + 1) * Q) * i32;
    bs[2 * 6 + 0 * 2 + 0] = (sa * (r - 1) * P - sb * (s - 1) * Q) * i16;
    bs[2 * 6 + 1 * 2 + 0] = (sa * (Y2 - Y3) * (r - 1) * P + sb * (Y3 - Y4) * (s - 1) * Q) * i32;
    bs[2 * 6 + 2 * 2 + 0] = (-sa * (X2 - X3) * (r - 1) * P - sb * (X3 - X4) * (s - 1) * Q) * i32;
    bs[2 * 6 + 0 * 2 + 1] = (-ca * (r - 1) * P + cb * (s - 1) * Q) * i16;
    bs[2 * 6 + 1 * 2 + 1] = (-ca * (Y2 - Y3) * (r - 1) * P - cb * (Y3 - Y4) * (s - 1) * Q) * i32;
    bs[2 * 6 + 2 * 2 + 1] = (ca * (X2 - X3) * (r - 1) * P + cb * (X3 - X4) * (s - 1) * Q) * i32;
    bs[3 * 6 + 0 * 2 + 0] = (-sa * (r + 1) * P + sb * (s - 1) * Q) * i16;
    bs[3 * 6 + 1 * 2 + 0] = (-sa * (Y1 - Y4) * (r + 1) * P + sb * (Y3 - Y4) * (s - 1) * Q) * i32;
    bs[3 * 6 + 2 * 2 + 0] = (sa * (X1 - X4) * (r + 1) * P - sb * (X3 - X4) * (s - 1) * Q) * i32;
    bs[3 * 6 + 0 * 2 + 1] = (ca * (r + 1) * P - cb * (s - 1) * Q) * i16;
    bs[3 * 6 + 1 * 2 + 1] = (ca * (Y1 - Y4) * (r + 1) * P - cb * (Y3 - Y4) * (s - 1) * Q) * i32;
    bs[3 * 6 + 2 * 2 + 1] = (-ca * (X1 - X4) * (r + 1) * P + cb * (X3 - X4) * (s - 1) * Q) * i32;
}


// ---------------------------------------------------------------------------
// CSR writes of node-pair blocks (scatter, fused tiles, boundary pairs)
// ---------------------------------------------------------------------------

// Row r of the upper pair (I, J): writes sum[0..5] (row dof(I, r), J's
// columns) and, for I != J, the same values as column dof(I, r) of the rows
// dof(J, c) -- K is bitwise symmetric by construction.
__device__ __forceinline__ void write_pair_row(const ScatterPlan &sp, int I, int J, int kI, int kJ, int r,
                                               const double sum[6], double *__restrict__ vals) {
    const int rsI = sp.node_rowstart[(int64_t)I * 6 + r];
    if (rsI >= 0) {
        const int8_t *codeJ = sp.node_code + (int64_t)J * 6;
        const int of = sp.nbr_off[2 * kI], ox = sp.nbr_off[2 * kI + 1];
        for (int c = 0; c < 6; ++c) {
            const int cd = codeJ[c];
            if (cd < 0) continue;
            vals[rsI + (cd < 8 ? of + cd : ox + cd - 8)] = sum[c];
        }
    }
    if (I != J) {
        const int cdI = sp.node_code[(int64_t)I * 6 + r];
        if (cdI >= 0) {
            const int pos = cdI < 8 ? sp.nbr_off[2 * kJ] + cdI : sp.nbr_off[2 * kJ + 1] + cdI - 8;
            for (int c = 0; c < 6; ++c) {
                const int rsJ = sp.node_rowstart[(int64_t)J * 6 + c];
                if (rsJ >= 0) vals[rsJ + pos] = sum[c];
            }
        }
    }
}

// Row r of a pair from its placement record (TilePlan trec / brec): the
// same writes as write_pair_row, with every position taken from the record.
// Codes of a node whose six dofs are all in the block, in order: the
// common case, written through one base address per run.
constexpr int FS_ALL_DOFS = 0x543210;

__device__ __forceinline__ void write_pair_row_rec(const int4 a, const int4 b, const int4 o, int r,
                                                   const double sum[6], double *__restrict__ vals) {
    const int crI = (a.w >> (4 * r)) & 15;
    if (crI == 15) return;
    const int rowI = crI < 8 ? a.x + crI * a.z : a.y + (crI - 8) * a.z;
    const bool allJ = (b.w & 0xffffff) == FS_ALL_DOFS;
    if (allJ) {
        double *d = vals + rowI + o.x;
        for (int c = 0; c < 6; ++c) d[c] = sum[c];
    } else {
        for (int c = 0; c < 6; ++c) {
            const int cc = (b.w >> (4 * c)) & 15;
            if (cc == 15) continue;
            vals[rowI + (cc < 8 ? o.x + cc : o.y + cc - 8)] = sum[c];
        }
    }
    if ((a.w >> 30) & 1) return;  // I = J: the block is its own transpose
    const int colI = crI < 8 ? o.z + crI : o.w + crI - 8;
    if (allJ) {
        double *d = vals + b.x + colI;
        for (int c = 0; c < 6; ++c) d[(int64_t)c * b.z] = sum[c];
    } else {
        for (int c = 0; c < 6; ++c) {
            const int cr = (b.w >> (4 * c)) & 15;
            if (cr == 15) continue;
            vals[(cr < 8 ? b.x + cr * b.z : b.y + (cr - 8) * b.z) + colI] = sum[c];
        }
    }
}

// Lane c of a pair from its placement record: col = column c of the pair's
// block K_IJ (col[r] = K_IJ[r][c]), row = row c (row[j] = K_IJ[c][j]).  The
// same values and positions as write_pair_row_rec over r = 0..5, grouped by
// column: value K_IJ[r][c] goes to row dof(I, r) at J's dof c, and for
// I != J, K_IJ[c][j] to row dof(J, j) at I's dof c.
__device__ __forceinline__ void write_pair_lane_rec(const int4 a, const int4 b, const int4 o, int c,
                                                    const double col[6], const double row[6],
                                                    double *__restrict__ vals) {
    const int ccJ = (b.w >> (4 * c)) & 15;
    if (ccJ != 15) {
        const int colJ = ccJ < 8 ? o.x + ccJ : o.y + ccJ - 8;
        for (int r = 0; r < 6; ++r) {
            const int crI = (a.w >> (4 * r)) & 15;
            if (crI == 15) continue;
            vals[(crI < 8 ? a.x + crI * a.z : a.y + (crI - 8) * a.z) + colJ] = col[r];
        }
    }
    if ((a.w >> 30) & 1) return;  // I = J: the block is its own transpose
    const int ccI = (a.w >> (4 * c)) & 15;
    if (ccI == 15) return;
    const int colI = ccI < 8 ? o.z + ccI : o.w + ccI - 8;
    for (int j = 0; j < 6; ++j) {
        const int cr = (b.w >> (4 * j)) & 15;
        if (cr == 15) continue;
        vals[(cr < 8 ? b.x + cr * b.z : b.y + (cr - 8) * b.z) + colI] = row[j];
    }
}

// Output phase of the fused assembly (TilePlan): thread (pair entry, row r)
// of tile t sums row r of the pair's blocks in the LDS stage (stage[(le NBLK
// + blk) 36 + 6 r + c], every block stored in the pair's orientation, rows
// of the lower-numbered node first) in element order, then writes the CSR
// values of a pair completed in the tile, or the 6 values of its partial
// sum.  A thread takes TW items per pass and
// loads all their records before the first sum, so that the record loads of
// the items overlap (C2 cylinder, fused Q4 kernel: 0.756 ms with one item,
// 0.727 ms with three; C4 plate, T3: 1.032 ms and 0.990 ms with four; the
// values are bitwise the same).
// Whether element e's block (a, b) is staged transposed: the fused output
// phase reads every block in the orientation of its node pair (I < J), the
// transposed bit of the host's contribution code.
__device__ __forceinline__ bool stage_transposed(int a, int b, int node_a, int node_b) {
    return a != b && node_a > node_b;
}

struct TileItem {
    int4 a, b, o;
    uint64_t cw;
    int slot;
};

__device__ __forceinline__ void tile_item_load(const TilePlan &tp, int q, TileItem &it) {
    it.cw = tp.tcontrib[q];
    it.slot = tp.tslot[q];
    it.a = tp.trec[3 * q];
    it.b = tp.trec[3 * q + 1];
    it.o = tp.trec[3 * q + 2];
}

template <int NBLK>
__device__ __forceinline__ void tile_item_write(const TileItem &it, int r, const TilePlan &tp, const double *stage,
                                                double *__restrict__ vals) {
    double sum[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int cnt = (int)(it.cw & 7);
    for (int k = 0; k < cnt; ++k) {
        const int v = (int)((it.cw >> (3 + 10 * k)) & 1023);
        const double *B = stage + ((v & 31) * NBLK + ((v >> 5) & 15)) * 36 + r * 6;
        for (int c = 0; c < 6; ++c) sum[c] += B[c];
    }
    if (it.slot >= 0) {
        double *dst = tp.partial + (int64_t)it.slot * 36 + r * 6;
        for (int c = 0; c < 6; ++c) dst[c] = sum[c];
    } else {
        write_pair_row_rec(it.a, it.b, it.o, r, sum, vals);
    }
}

template <int NBLK, int WG, int TW>
__device__ __forceinline__ void tile_write(int t, const TilePlan &tp, const double *stage,
                                           double *__restrict__ vals) {
    const int qb = tp.tp_ptr[t], qe = tp.tp_ptr[t + 1];
    const int n = (qe - qb) * 6;
    for (int i0 = threadIdx.x; i0 < n; i0 += TW * WG) {
        TileItem it[TW];
        for (int w = 0; w < TW; ++w) {
            const int i = i0 + w * WG;
            if (i < n) tile_item_load(tp, qb + i / 6, it[w]);
        }
        for (int w = 0; w < TW; ++w) {
            const int i = i0 + w * WG;
            if (i < n) tile_item_write<NBLK>(it[w], i % 6, tp, stage, vals);
        }
    }
}

template <int NEN, int NQP, int NR, int NE, int WG, int SPLIT, int RS, bool FUSED, bool COMP>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(2))) void elem_kernel(ElemParams p, const double *__restrict__ xs,
                                                  const double *__restrict__ ys,
                                                  const double *__restrict__ zs,
                                                  const int32_t *__restrict__ conn,
                                                  const double *__restrict__ nxs,
                                                  const double *__restrict__ nys,
                                                  const double *__restrict__ nzs,
                                                  const uint8_t *__restrict__ valid,
                                                  double *__restrict__ Kb, int *__restrict__ status,
                                                  ScatterPlan sp, TilePlan tp, double *__restrict__ vals) {
    using T = ElemTraits<NEN, NQP, NR>;
    using S = ElemSmem<NEN, NQP, NR, NE>;
    static_assert(!FUSED || SPLIT == 0, "the fused assembly writes whole blocks");
    extern __shared__ double smem[];
    double *sC = smem + S::OFF_C;
    double *sG = smem + S::OFF_G;
    double *sN = smem + S::OFF_N;
    double *sM = smem + S::OFF_M;

    const int tid = threadIdx.x;
    const int E = p.nelem;
    // fused: workgroups visit the tiles XCD by XCD, so that the neighbouring
    // tiles writing adjacent stretches of the same CSR rows share an L2
    const int tblk = FUSED ? (int)xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int e0 = p.ebegin + tblk * NE;
    const int ne = FUSED ? NE : ((p.eend - e0) < NE ? (p.eend - e0) : NE);
    auto elem = [&](int le) -> int { return FUSED ? p.tile_elem[e0 + le] : e0 + le; };
    // node a of tile element le: the fused tiles carry their own copy of the
    // connectivity, so coordinate and normal loads wait on one index load
    auto enode = [&](int le, int a) -> int { return FUSED ? p.tile_conn[(e0 + le) * NEN + a] : conn[a * E + e0 + le]; };
    // node data of the thread's G2 lane (element, node), loaded before G1 so
    // that the loads overlap it
    uint8_t valid2 = 0;
    double n20 = 0.0, n21 = 0.0, n22 = 0.0;
    {
        const int le = tid / (NQP * NEN * RS), j = (tid / RS) % NEN;
        if (tid < NE * NQP * NEN * RS && le < ne) {
            const int node = enode(le, j);
            valid2 = valid[node];
            n20 = nxs[node];
            n21 = nys[node];
            n22 = nzs[node];
        }
    }
    // block lane (element, node pair) of the C phase
    const bool cthread = tid < NE * T::NBLK;
    const int le3 = tid / T::NBLK;
    const int blk = tid % T::NBLK;
    int ba = 0, bb = 0;
    {
        int cnt = 0;
        for (int a = 0; a < NEN; ++a)
            for (int b = a; b < NEN; ++b) {
                if (cnt == blk) {
                    ba = a;
                    bb = b;
                }
                ++cnt;
            }
    }
    const bool cact = cthread && (le3 < ne);
    // fused: the block lane's node data, loaded now as well: the flag of
    // node a (the drilling term of a diagonal block) and the staged
    // orientation (the two-kernel form measured faster loading the flag late)
    bool cvalid = false, ctrans = false;
    if (FUSED && cact) {
        const int na = enode(le3, ba), nb = enode(le3, bb);
        cvalid = valid[na];
        ctrans = stage_transposed(ba, bb, na, nb);
    }

    // ---------------- G1: geometry per element (FEMMShellT3FFModule.jl:269-321,465-484) ----------------
    // One thread per element; each part of the record (frame and gradients,
    // material factors, DSG shear entries) is stored as soon as it is
    // computed, which keeps the thread's live registers (and the kernel's) low.
    static_assert(NEN == 3 && NQP == 1, "elem_kernel is the T3 kernel");
    if (tid < NE && tid < ne) {
        const int le = tid;
        const int e = elem(le);
        double X[NEN][3];
        for (int a = 0; a < NEN; ++a) {
            const int n = enode(le, a);
            X[a][0] = xs[n];
            X[a][1] = ys[n];
            X[a][2] = zs[n];
        }
        double *g = sG + le * T::GEO;
        double EG[3][3];
        double J0[3][2];
        for (int j = 0; j < 3; ++j) {
            J0[j][0] = X[1][j] - X[0][j];
            J0[j][1] = X[2][j] - X[0][j];
        }
        e_g(EG, J0);
        double ec[3][2];
        ec[0][0] = 0.0;
        ec[0][1] = 0.0;
        ec[1][0] = J0[0][0] * EG[0][0] + J0[1][0] * EG[1][0] + J0[2][0] * EG[2][0];
        ec[1][1] = J0[0][0] * EG[0][1] + J0[1][0] * EG[1][1] + J0[2][0] * EG[2][1];
        ec[2][0] = J0[0][1] * EG[0][0] + J0[1][1] * EG[1][0] + J0[2][1] * EG[2][0];
        ec[2][1] = J0[0][1] * EG[0][1] + J0[1][1] * EG[1][1] + J0[2][1] * EG[2][1];
        const double a = ec[1][0], b = ec[1][1], c = ec[2][0], d = ec[2][1];
        const double Jd = (a * d - b * c);
        const double Ae = Jd / 2;
        {
            double gN[NEN][2];
            gN[0][0] = (b - d) / Jd;
            gN[1][0] = d / Jd;
            gN[2][0] = -b / Jd;
            gN[0][1] = (c - a) / Jd;
            gN[1][1] = -c / Jd;
            gN[2][1] = a / Jd;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) g[T::G_EG + i * 3 + j] = EG[i][j];
            for (int k = 0; k < NEN; ++k) {
                g[T::G_GN + 2 * k] = gN[k][0];
                g[T::G_GN + 2 * k + 1] = gN[k][1];
            }
        }
        {
            double Lmb[36], Ls[4];
            const int grp = p.elem_group ? p.elem_group[e] : 0;
            const double *hrec = COMP ? nullptr : p.g_homog + 16 * grp;
            const double tt = COMP ? p.g_t[grp] : hrec[13];
            const double stab_c = COMP ? p.stab_c : hrec[14];
            const double h = sqrt(2 * Ae);
            const double stab = (tt * tt) / (tt * tt + stab_c * (h * h));
            const double wsh = (NR == 12) ? Ae / 3 : Ae;  // AVERAGE_K splits the weight
            if constexpr (!COMP) {
                homog_factors(hrec, hrec + 9, sqrt(tt * Ae), sqrt((tt * tt * tt) / 12 * Ae),
                              sqrt(tt * stab * wsh), Lmb, Ls);
            } else {
                double m, n;
                layup_angle(EG, p.g_csys + 9 * grp, m, n);
                composite_factors(p.g_Labd + 36 * grp, p.g_Lh + 4 * grp, m, n, sqrt(Ae), sqrt(stab * wsh),
                                  Lmb, Ls);
            }
            for (int i = 0; i < 36; ++i) g[T::G_LMB + i] = Lmb[i];
            for (int i = 0; i < 4; ++i) g[T::G_LS + i] = Ls[i];
        }
        {
            // DSG transverse shear entries (_add_Bsmat_o!, :486-537)
            double bs[3][3][T::NSH];
            for (int j = 0; j < 3; ++j)
                for (int cc = 0; cc < 3; ++cc)
                    for (int r = 0; r < T::NSH; ++r) bs[j][cc][r] = 0.0;
            const int ord[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
            for (int o = 0; o < 3; ++o) {
                const int s = ord[o][0], pp = ord[o][1], qq = ord[o][2];
                const double aa = ec[pp][0] - ec[s][0];
                const double bb = ec[pp][1] - ec[s][1];
                const double cc = ec[qq][0] - ec[s][0];
                const double dd = ec[qq][1] - ec[s][1];
                const double mm = (1.0 / 2 / Ae);
                const int r0 = (NR == 12) ? 2 * o : 0;
                // comps: 0 = w, 1 = theta_x, 2 = theta_y
                bs[s][0][r0] += mm * (bb - dd);
                bs[s][2][r0] += mm * (Ae);
                bs[s][0][r0 + 1] += mm * (cc - aa);
                bs[s][1][r0 + 1] += mm * (-Ae);
                bs[pp][0][r0] += mm * (dd);
                bs[pp][1][r0] += mm * (-bb * dd / 2);
                bs[pp][2][r0] += mm * (aa * dd / 2);
                bs[pp][0][r0 + 1] += mm * (-cc);
                bs[pp][1][r0 + 1] += mm * (bb * cc / 2);
                bs[pp][2][r0 + 1] += mm * (-aa * cc / 2);
                bs[qq][0][r0] += mm * (-bb);
                bs[qq][1][r0] += mm * (bb * dd / 2);
                bs[qq][2][r0] += mm * (-bb * cc / 2);
                bs[qq][0][r0 + 1] += mm * (aa);
                bs[qq][1][r0 + 1] += mm * (-aa * dd / 2);
                bs[qq][2][r0 + 1] += mm * (aa * cc / 2);
            }
            if (NR == 8) {
                for (int j = 0; j < 3; ++j)
                    for (int cc = 0; cc < 3; ++cc)
                        for (int r = 0; r < 2; ++r) bs[j][cc][r] *= (1.0 / 3);
            }
            for (int j = 0; j < 3; ++j)
                for (int cc = 0; cc < 3; ++cc)
                    for (int r = 0; r < T::NSH; ++r) g[T::G_BS + (j * 3 + cc) * T::NSH + r] = bs[j][cc][r];
        }
    }
    __syncthreads();

    // ---------------- G2/G3: per (element, qp, node) ----------------
    // Rows of C are produced one at a time from the material factor rows in
    // the geometry record, so only the triad stays in registers across the
    // barrier.
    // Lane (element, qp, node, half): with RS = 2 the two halves of a node
    // take rows 3h..3h+2 of the membrane/bending factor and half of the
    // shear rows each (the same control flow on both halves).
    static_assert(RS == 1 || RS == 2, "row split");
    constexpr int RPL = NR / RS;  // rows of C per lane
    const bool gthread = tid < NE * NQP * NEN * RS;
    const int le2 = tid / (NQP * NEN * RS);
    const int q2 = (tid / (NEN * RS)) % NQP;
    const int j2 = (tid / RS) % NEN;
    const int h2 = tid % RS;
    const bool gact = gthread && (le2 < ne);
    auto row_of = [&](int i) { return RS == 1 ? i : (i < 3 ? 3 * h2 + i : 6 + h2 * (T::NSH / 2) + (i - 3)); };
    double A[3][3];
    if (gact) {
        const double *g = sG + (le2 * NQP + q2) * T::GEO;
        double nk_e[3];
        if (valid2) {
            for (int i = 0; i < 3; ++i)
                nk_e[i] = g[T::G_EG + 0 * 3 + i] * n20 + g[T::G_EG + 1 * 3 + i] * n21 + g[T::G_EG + 2 * 3 + i] * n22;
        } else {
            nk_e[0] = 0.0;
            nk_e[1] = 0.0;
            nk_e[2] = 1.0;
        }
        nodal_triad(nk_e, A);
        const double gNj0 = g[T::G_GN + 2 * j2], gNj1 = g[T::G_GN + 2 * j2 + 1];
        const double *bs = g + T::G_BS + j2 * 3 * T::NSH;
        // coupling vector d = m1 * cx + m2 * cy (_transfmat_a_to_e!, :451-460)
        const double m1 = (1 / A[2][2]) * A[0][2];
        const double m2 = (1 / A[2][2]) * A[1][2];
        double *nd = sN + ((le2 * NQP + q2) * NEN + j2) * T::NODE;
        if (h2 == 0)
            for (int k = 0; k < 3; ++k) {
                nd[k] = A[0][k];
                nd[3 + k] = A[1][k];
            }
        for (int i = 0; i < RPL; ++i) {
            const int s = row_of(i);
            double cu, cv, cw, cx, cy;
            if (i < 3 || (RS == 1 && s < 6))
                c_row_mb(g + T::G_LMB + s * 6, gNj0, gNj1, cu, cv, cw, cx, cy);
            else
                c_row_sh<T::NSH>(g + T::G_LS, s, bs, cu, cv, cw, cx, cy);
            nd[6 + s] = m1 * cx + m2 * cy;
        }
    }
    __syncthreads();
    if (gact) {
        const double *g = sG + (le2 * NQP + q2) * T::GEO;
        const double *ndb = sN + ((le2 * NQP + q2) * NEN) * T::NODE;
        const double gNj0 = g[T::G_GN + 2 * j2], gNj1 = g[T::G_GN + 2 * j2 + 1];
        const double *bs = g + T::G_BS + j2 * 3 * T::NSH;
        // Q = A' * EG' : global -> nodal block (_TransfmatGToA, :398-419)
        double Q[3][3];
        for (int k = 0; k < 3; ++k)
            for (int gg = 0; gg < 3; ++gg)
                Q[k][gg] = A[0][k] * g[T::G_EG + gg * 3 + 0] + A[1][k] * g[T::G_EG + gg * 3 + 1] +
                           A[2][k] * g[T::G_EG + gg * 3 + 2];
        // translations: cn_k = sum_rw c_rw A[rw][k] + sum_i a3(i,j,k) d_i
        double a3[3][NEN];
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < NEN; ++i)
                a3[k][i] = 1.0 / 2 * (ndb[i * T::NODE + 3 + k] * gNj0 - ndb[i * T::NODE + k] * gNj1);
        // rotations: nodal theta_1,2 from element theta_x,y; drilling column zero
        double r2a[2], r2b[2];
        for (int l = 0; l < 2; ++l) {
            r2a[l] = A[0][l] - (1 / A[2][2]) * A[0][2] * A[l][2];
            r2b[l] = A[1][l] - (1 / A[2][2]) * A[1][2] * A[l][2];
        }
        double kp[2] = {0.0, 0.0};
        double *cdst = sC + le2 * T::CSTR + (q2 * NR) * T::K + 6 * j2;
        for (int i = 0; i < RPL; ++i) {
            const int s = row_of(i);
            double cu, cv, cw, cx, cy;
            if (i < 3 || (RS == 1 && s < 6))
                c_row_mb(g + T::G_LMB + s * 6, gNj0, gNj1, cu, cv, cw, cx, cy);
            else
                c_row_sh<T::NSH>(g + T::G_LS, s, bs, cu, cv, cw, cx, cy);
            double cgt[3] = {0.0, 0.0, 0.0}, cgr[3] = {0.0, 0.0, 0.0};
            for (int k = 0; k < 3; ++k) {
                double cn = cu * A[0][k] + cv * A[1][k] + cw * A[2][k];
                for (int i = 0; i < NEN; ++i) cn += a3[k][i] * ndb[i * T::NODE + 6 + s];
                for (int gg = 0; gg < 3; ++gg) cgt[gg] += cn * Q[k][gg];
            }
            for (int l = 0; l < 2; ++l) {
                const double cn = cx * r2a[l] + cy * r2b[l];
                kp[l] += cn * cn;
                for (int gg = 0; gg < 3; ++gg) cgr[gg] += cn * Q[l][gg];
            }
            for (int gg = 0; gg < 3; ++gg) {
                cdst[s * T::K + gg] = cgt[gg];
                cdst[s * T::K + 3 + gg] = cgr[gg];
            }
        }
        if (RS == 2)  // both halves of the node: half 0's sum + half 1's sum
            for (int l = 0; l < 2; ++l) {
                const double other = __shfl_xor(kp[l], 1);
                kp[l] = h2 == 0 ? kp[l] + other : other + kp[l];
            }
        if (NQP == 1 && h2 == 0) {
            double *ms = sM + le2 * T::MISC;
            for (int gg = 0; gg < 3; ++gg) ms[T::M_G + 3 * j2 + gg] = Q[2][gg];
            ms[T::M_KP + 2 * j2] = kp[0];
            ms[T::M_KP + 2 * j2 + 1] = kp[1];
        }
    }
    __syncthreads();

    // ---------------- C: node-pair blocks ----------------
    constexpr bool split_sh = SPLIT == 1;  // shear rows leave as factors
    constexpr bool split_dr = SPLIT != 0;  // drilling blocks leave separately
    double acc[6][6];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) acc[r][c] = 0.0;
    if (cact) {
        const double *cb = sC + le3 * T::CSTR;
        // Homogeneous shells with one quadrature point: rows 0..2 of C are
        // the membrane rows, whose factor rows have exact zeros in the
        // bending columns (Lmb is block diagonal), so cx = cy = 0 and their
        // rotation columns are signed zeros.  Their products with those
        // columns add signed zeros to +0 or to nonzero sums: leaving them out
        // keeps every block bitwise the same.
        constexpr bool memb3 = NQP == 1 && !COMP;
        if (memb3) {
            for (int s = 0; s < 3; ++s) {
                double va[3], vb[3];
                for (int r = 0; r < 3; ++r) {
                    va[r] = cb[s * T::K + 6 * ba + r];
                    vb[r] = cb[s * T::K + 6 * bb + r];
                }
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) acc[r][c] += va[r] * vb[c];
            }
        }
        for (int s = memb3 ? 3 : 0; s < T::NS; ++s) {
            if (split_sh && (s % NR) >= 6) continue;
            double va[6], vb[6];
            for (int r = 0; r < 6; ++r) {
                va[r] = cb[s * T::K + 6 * ba + r];
                vb[r] = cb[s * T::K + 6 * bb + r];
            }
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) acc[r][c] += va[r] * vb[c];
        }
        if (NEN == 3 && ba == bb) {
            // T3 drilling in the nodal basis (:714-727) as a global rank-one term
            const int e = elem(le3);
            if (!FUSED) cvalid = valid[enode(le3, ba)];
            const double *ms = sM + le3 * T::MISC;
            const int grp = p.elem_group ? p.elem_group[e] : 0;
            const double dscale = COMP ? p.drill_scale : p.g_homog[16 * grp + 15];
            const double kavg = (((((ms[T::M_KP + 0] + ms[T::M_KP + 2]) + ms[T::M_KP + 4]) + ms[T::M_KP + 1]) +
                                   ms[T::M_KP + 3]) + ms[T::M_KP + 5]) / 6 * dscale;
            const double *gv = ms + T::M_G + 3 * ba;
            if (split_dr) {
                double *dd = p.split_dr + ((size_t)e * NEN + ba) * 9;
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) dd[r * 3 + c] = cvalid ? kavg * (gv[r] * gv[c]) : 0.0;
            } else if (cvalid) {
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) acc[3 + r][3 + c] += kavg * (gv[r] * gv[c]);
            }
        }
    }
    if (split_sh) {
        // shear rows of C (global basis) of every element of the block
        constexpr int NSR = NQP * T::NSH;
        for (int i = tid; i < ne * NSR * T::K; i += WG) {
            const int le = i / (NSR * T::K), rem = i % (NSR * T::K);
            const int row = rem / T::K, col = rem % T::K;
            const int srow = (row / T::NSH) * NR + 6 + row % T::NSH;
            p.split_cs[(size_t)(e0 + le) * NSR * T::K + rem] = sC[le * T::CSTR + srow * T::K + col];
        }
    }
    __syncthreads();  // every read of sC and sG is done: reuse them as the output stage
    if (cact) {
        double *ks = sC + (le3 * T::NBLK + blk) * 36;
        if (FUSED && ctrans) {
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) ks[c * 6 + r] = acc[r][c];
        } else {
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) ks[r * 6 + c] = acc[r][c];
        }
    }
    __syncthreads();
    if constexpr (FUSED) {
        tile_write<T::NBLK, WG, 4>(tblk, tp, sC, vals);  // 4 items per pass: measured best of 1-4 on T3
    } else {
        const int total = ne * T::KSTAGE;  // contiguous in Kb
        double *dst = Kb + (size_t)e0 * T::KSTAGE;
        for (int i = tid; i < total; i += WG) dst[i] = sC[i];
    }
}

// ---------------------------------------------------------------------------
// Q4 element kernel, Gauss points processed in sequence
// ---------------------------------------------------------------------------
//
// The phases of the T3 kernel (G1 geometry, G2/G3 rows of C, C blocks),
// with the LDS holding the C matrix of one Gauss point at a time (8 x 24):
// the K blocks accumulate in registers of the (element, node pair) threads
// across the four Gauss points, so an element needs ~3.7 KB of LDS where all
// four points of C would take ~11 KB.  Phase G1 (per element and Gauss
// point) stores a compact geometry record: E_G, gradN_e, the material scale
// factors (and the layup angle for layered shells) and the MITC shear row
// entries; the material factors are rebuilt per node thread.
template <int NE, int WG, int SPLIT, int RS, bool COMP, bool FUSED, bool UG>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(2))) void elem_q4_seq_kernel(ElemParams p, const double *__restrict__ xs,
                                                         const double *__restrict__ ys,
                                                         const double *__restrict__ zs,
                                                         const int32_t *__restrict__ conn,
                                                         const double *__restrict__ nxs,
                                                         const double *__restrict__ nys,
                                                         const double *__restrict__ nzs,
                                                         const uint8_t *__restrict__ valid,
                                                         double *__restrict__ Kb, int *__restrict__ status,
                                                         ScatterPlan sp, TilePlan tp, double *__restrict__ vals) {
    constexpr int NEN = 4, NQP = 4, NR = 8, K = 24, NBLK = 10;
    static_assert(!FUSED || SPLIT == 0, "the fused assembly writes whole blocks");
    // geometry record per (element, Gauss point): EG 9 | gN 8 | s_m s_b s_s
    // (layered: s_mb 0 s_s m n) | bs 24 (layered: +1 pad)
    constexpr int NSC = COMP ? 5 : 3;
    constexpr int GEO = 17 + NSC + 24 + (COMP ? 1 : 0);
    constexpr int G_EG = 0, G_GN = 9, G_SC = 17, G_BS = 17 + NSC;
    constexpr int CSTR = NR * K + 2;    // one Gauss point of C, padded
    constexpr int NODE = 6 + NR;
    // C first: the accumulation's LDS reads then address row s of C at
    // immediate offsets (< 2 KB) from one base per node
    constexpr int OFF_C = 0;
    constexpr int OFF_G = OFF_C + NE * CSTR;
    constexpr int OFF_N = OFF_G + NE * NQP * GEO;
    // after the Gauss-point loop the node records are dead: the drilling
    // scratch (rotation blocks K_rr of the 4 nodes | trace terms) takes their place
    constexpr int OFF_M = OFF_N;
    constexpr int MISC = 4 * 9 + 4 + 1;
    static_assert(MISC <= NEN * NODE, "drilling scratch must fit in the node records");
    static_assert(NE * NQP * GEO + NE * CSTR >= NE * NBLK * 36, "output stage must fit in the G+C region");
    extern __shared__ double smem[];
    double *sG = smem + OFF_G;
    double *sC = smem + OFF_C;
    double *sN = smem + OFF_N;
    double *sM = smem + OFF_M;

    const int tid = threadIdx.x;
    const int E = p.nelem;
    const int tblk = FUSED ? (int)xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int e0 = p.ebegin + tblk * NE;
    const int ne = FUSED ? NE : ((p.eend - e0) < NE ? (p.eend - e0) : NE);
    auto elem = [&](int le) -> int { return FUSED ? p.tile_elem[e0 + le] : e0 + le; };
    // node a of tile element le: the fused tiles carry their own copy of the
    // connectivity, so coordinate and normal loads wait on one index load
    auto enode = [&](int le, int a) -> int { return FUSED ? p.tile_conn[(e0 + le) * NEN + a] : conn[a * E + e0 + le]; };

    // thread roles: G2/G3 lane (element, node, row half) -- RS lanes share
    // the eight rows of C of a node
    static_assert(RS == 1 || RS == 2, "row split");
    const bool gact = tid < NE * NEN * RS && (tid / (NEN * RS)) < ne;
    const int le2 = tid / (NEN * RS), j2 = (tid / RS) % NEN, h2 = tid % RS;
    int node2 = 0;
    uint8_t valid2 = 0;
    double n20 = 0.0, n21 = 0.0, n22 = 0.0;
    // material group of the lane's element: with one group (UG) a constant,
    // so the material factor rows come through scalar loads of one record
    int grp2 = 0;
    if (!UG && gact && p.elem_group) grp2 = p.elem_group[elem(le2)];
    if (gact) {  // loaded before G1 so that the loads overlap it
        node2 = enode(le2, j2);
        valid2 = valid[node2];
        n20 = nxs[node2];
        n21 = nys[node2];
        n22 = nzs[node2];
    }
    // block lane (element, node pair) of the accumulation, and its staged
    // orientation, loaded now too
    const bool cact = tid < NE * NBLK && (tid / NBLK) < ne;
    const int le3 = tid / NBLK, blk = tid % NBLK;
    int ba = 0, bb = 0;
    {
        int cnt = 0;
        for (int a = 0; a < NEN; ++a)
            for (int b = a; b < NEN; ++b) {
                if (cnt == blk) {
                    ba = a;
                    bb = b;
                }
                ++cnt;
            }
    }
    const bool ctrans = FUSED && cact && stage_transposed(ba, bb, enode(le3, ba), enode(le3, bb));
    // ---------------- G1: geometry per (element, Gauss point) ----------------
    // With two or more waves, wave 0 computes the frame, gradients and
    // material scales of every (element, Gauss point) of the tile and wave 1
    // its MITC shear entries (both from the same tangents and basis); the
    // branches are uniform per wave.  A single wave does both.
    constexpr int NW = WG / 64;
    constexpr bool G1_TWO = NW >= 2 && NE * NQP <= 64;
    const int g1_roles = G1_TWO ? ((tid >> 6) == 0 ? 1 : (tid >> 6) == 1 ? 2 : 0) : 3;
    const int g1_rec = G1_TWO ? (tid & 63) : tid;
    if (g1_roles != 0 && g1_rec < NE * NQP) {
        const int le = g1_rec / NQP, q = g1_rec % NQP;
        if (le < ne) {
            const int e = elem(le);
            double X[4][3];
            for (int a = 0; a < 4; ++a) {
                const int n = enode(le, a);
                X[a][0] = xs[n];
                X[a][1] = ys[n];
                X[a][2] = zs[n];
            }
            double *g = sG + (le * NQP + q) * GEO;
            const double xi = (q / 2 == 0) ? -FS_GAUSS2 : FS_GAUSS2;
            const double eta = (q % 2 == 0) ? -FS_GAUSS2 : FS_GAUSS2;
            double gNp[4][2];
            q4_gradn(xi, eta, gNp);
            double J[3][2];
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 2; ++k) {
                    double acc = 0.0;
                    for (int a = 0; a < 4; ++a) acc += X[a][i] * gNp[a][k];
                    J[i][k] = acc;
                }
            double EG[3][3];
            e_g(EG, J);
            if (g1_roles & 1) {
                const int grp = UG ? 0 : (p.elem_group ? p.elem_group[e] : 0);
                const double *hrec = COMP ? nullptr : p.g_homog + 16 * grp;
                const double tt = COMP ? p.g_t[grp] : hrec[13];
                const double stab_c = COMP ? p.stab_c : hrec[14];
                const double cr0 = -J[2][0] * J[1][1] + J[1][0] * J[2][1];
                const double cr1 = J[2][0] * J[0][1] - J[0][0] * J[2][1];
                const double cr2 = -J[1][0] * J[0][1] + J[0][0] * J[1][1];
                const double Jac = sqrt(cr0 * cr0 + cr1 * cr1 + cr2 * cr2);
                const double G00 = J[0][0] * J[0][0] + J[1][0] * J[1][0] + J[2][0] * J[2][0];
                const double G01 = J[0][0] * J[0][1] + J[1][0] * J[1][1] + J[2][0] * J[2][1];
                const double G11 = J[0][1] * J[0][1] + J[1][1] * J[1][1] + J[2][1] * J[2][1];
                const double detG = G00 * G11 - G01 * G01;
                if (detG == 0.0) status[0] = 1;
                const double idetG = 1.0 / detG;
                const double i00 = G11 * idetG, i01 = -G01 * idetG, i11 = G00 * idetG;
                for (int a = 0; a < 4; ++a) {
                    const double t0 = i00 * gNp[a][0] + i01 * gNp[a][1];
                    const double t1 = i01 * gNp[a][0] + i11 * gNp[a][1];
                    double gg[3];
                    for (int i = 0; i < 3; ++i) gg[i] = J[i][0] * t0 + J[i][1] * t1;
                    g[G_GN + 2 * a] = EG[0][0] * gg[0] + EG[1][0] * gg[1] + EG[2][0] * gg[2];
                    g[G_GN + 2 * a + 1] = EG[0][1] * gg[0] + EG[1][1] * gg[1] + EG[2][1] * gg[2];
                }
                double maxd = 0.0;
                for (int i = 1; i < 4; ++i) {
                    const double dx = X[i][0] - X[0][0], dy = X[i][1] - X[0][1], dz = X[i][2] - X[0][2];
                    const double dist = (dx * dx + dy * dy + dz * dz);
                    maxd = dist > maxd ? dist : maxd;
                }
                const double h = sqrt(maxd);
                const double stab = (tt * tt) / (tt * tt + stab_c * (h * h));
                if (!COMP) {
                    g[G_SC + 0] = sqrt(tt * Jac);
                    g[G_SC + 1] = sqrt((tt * tt * tt / 12.0) * Jac);
                    g[G_SC + 2] = sqrt(tt * stab * Jac);
                } else {
                    double m, n;
                    layup_angle(EG, p.g_csys + 9 * grp, m, n);
                    g[G_SC + 0] = sqrt(Jac);
                    g[G_SC + 1] = 0.0;
                    g[G_SC + 2] = sqrt(stab * Jac);
                    g[G_SC + 3] = m;
                    g[G_SC + 4] = n;
                }
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) g[G_EG + i * 3 + j] = EG[i][j];
            }
            if (g1_roles & 2) {
                double cen[3] = {0.0, 0.0, 0.0};
                for (int a = 0; a < 4; ++a)
                    for (int k = 0; k < 3; ++k) cen[k] += X[a][k];
                for (int k = 0; k < 3; ++k) cen[k] /= 4;
                double ec[4][2];
                for (int a = 0; a < 4; ++a)
                    for (int k = 0; k < 2; ++k) {
                        double dd = 0.0;
                        for (int m = 0; m < 3; ++m) dd += (X[a][m] - cen[m]) * EG[m][k];
                        ec[a][k] = dd;
                    }
                q4_mitc_bs(ec, xi, eta, g + G_BS);
            }
        }
    }

    constexpr bool split_sh = SPLIT == 1;  // shear rows leave as factors
    constexpr bool split_dr = SPLIT != 0;  // drilling blocks leave separately
    double acc[6][6];
    double accs[3][3];  // split_sh: shear part of the rotation block (drilling penalty input)
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) acc[r][c] = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) accs[r][c] = 0.0;
    __syncthreads();

    for (int q = 0; q < NQP; ++q) {
        // ---------------- G2: triad and coupling vector of node j ----------------
        // Rows of C are produced one at a time (element-frame columns from the
        // material factor row, then the global-basis columns), so only the
        // triad and a few scalars stay in registers across the phases.
        double A[3][3];
        const double *g = sG + (le2 * NQP + q) * GEO;
        if (gact) {
            double nk_e[3];
            if (valid2) {
                for (int i = 0; i < 3; ++i)
                    nk_e[i] = g[G_EG + 0 * 3 + i] * n20 + g[G_EG + 1 * 3 + i] * n21 + g[G_EG + 2 * 3 + i] * n22;
            } else {
                nk_e[0] = 0.0;
                nk_e[1] = 0.0;
                nk_e[2] = 1.0;
            }
            nodal_triad(nk_e, A);
            const MatRows mr = mat_rows(p, COMP, grp2, g[G_SC + 0], g[G_SC + 1], g[G_SC + 2],
                                        COMP ? g[G_SC + NSC - 2] : 0.0, COMP ? g[G_SC + NSC - 1] : 0.0);
            const double gNj0 = g[G_GN + 2 * j2], gNj1 = g[G_GN + 2 * j2 + 1];
            const double *bs = g + G_BS + j2 * 6;
            const double m1 = (1 / A[2][2]) * A[0][2];
            const double m2 = (1 / A[2][2]) * A[1][2];
            double *nd = sN + (le2 * NEN + j2) * NODE;
            if (h2 == 0)
                for (int k = 0; k < 3; ++k) {
                    nd[k] = A[0][k];
                    nd[3 + k] = A[1][k];
                }
            if (RS == 1) {
                for (int s = 0; s < NR; ++s) {
                    double cu, cv, cw, cx, cy;
                    c_row_q4(mr, s, gNj0, gNj1, bs, cu, cv, cw, cx, cy);
                    nd[6 + s] = m1 * cx + m2 * cy;
                }
            } else {
                for (int i = 0; i < NR / 2; ++i) {
                    double cu, cv, cw, cx, cy;
                    c_row_q4_split(mr, i, h2, gNj0, gNj1, bs, cu, cv, cw, cx, cy);
                    nd[6 + q4_split_row(i, h2)] = m1 * cx + m2 * cy;
                }
            }
        }
        __syncthreads();
        // ---------------- G3: global-basis C columns of node j ----------------
        if (gact) {
            const double *ndb = sN + (le2 * NEN) * NODE;
            const double gNj0 = g[G_GN + 2 * j2], gNj1 = g[G_GN + 2 * j2 + 1];
            double Q[3][3];
            for (int k = 0; k < 3; ++k)
                for (int gg = 0; gg < 3; ++gg)
                    Q[k][gg] = A[0][k] * g[G_EG + gg * 3 + 0] + A[1][k] * g[G_EG + gg * 3 + 1] +
                               A[2][k] * g[G_EG + gg * 3 + 2];
            double a3[3][NEN];
            for (int k = 0; k < 3; ++k)
                for (int i = 0; i < NEN; ++i)
                    a3[k][i] = 1.0 / 2 * (ndb[i * NODE + 3 + k] * gNj0 - ndb[i * NODE + k] * gNj1);
            double r2a[2], r2b[2];
            for (int l = 0; l < 2; ++l) {
                r2a[l] = A[0][l] - (1 / A[2][2]) * A[0][2] * A[l][2];
                r2b[l] = A[1][l] - (1 / A[2][2]) * A[1][2] * A[l][2];
            }
            const MatRows mr = mat_rows(p, COMP, grp2, g[G_SC + 0], g[G_SC + 1], g[G_SC + 2],
                                        COMP ? g[G_SC + NSC - 2] : 0.0, COMP ? g[G_SC + NSC - 1] : 0.0);
            const double *bs = g + G_BS + j2 * 6;
            double *cdst = sC + le2 * CSTR + 6 * j2;
            // global-basis columns of node j for row s of C
            auto emit = [&](int s, double cu, double cv, double cw, double cx, double cy) {
                double cgt[3] = {0.0, 0.0, 0.0}, cgr[3] = {0.0, 0.0, 0.0};
                for (int k = 0; k < 3; ++k) {
                    double cn = cu * A[0][k] + cv * A[1][k] + cw * A[2][k];
                    for (int i = 0; i < NEN; ++i) cn += a3[k][i] * ndb[i * NODE + 6 + s];
                    for (int gg = 0; gg < 3; ++gg) cgt[gg] += cn * Q[k][gg];
                }
                for (int l = 0; l < 2; ++l) {
                    const double cn = cx * r2a[l] + cy * r2b[l];
                    for (int gg = 0; gg < 3; ++gg) cgr[gg] += cn * Q[l][gg];
                }
                for (int gg = 0; gg < 3; ++gg) {
                    cdst[s * K + gg] = cgt[gg];
                    cdst[s * K + 3 + gg] = cgr[gg];
                }
            };
            if (RS == 1) {
                for (int s = 0; s < NR; ++s) {
                    double cu, cv, cw, cx, cy;
                    c_row_q4(mr, s, gNj0, gNj1, bs, cu, cv, cw, cx, cy);
                    emit(s, cu, cv, cw, cx, cy);
                }
            } else {
                for (int i = 0; i < NR / 2; ++i) {
                    double cu, cv, cw, cx, cy;
                    c_row_q4_split(mr, i, h2, gNj0, gNj1, bs, cu, cv, cw, cx, cy);
                    emit(q4_split_row(i, h2), cu, cv, cw, cx, cy);
                }
            }
        }
        __syncthreads();
        // ---------------- C: accumulate node-pair blocks ----------------
        if (cact) {
            const double *cbp = sC + le3 * CSTR;
            // Homogeneous shells: rows 0..2 of C are the membrane rows (with
            // either row split), which have zero rotation columns (c_row_mb
            // with lrow[3..5] = 0 gives cx = cy = 0, so no coupling term and
            // cgr = 0).  Their products with the rotation columns are signed
            // zeros added to +0 or to nonzero sums, so leaving them out keeps
            // every block bitwise the same.
            if (!COMP) {
                for (int s = 0; s < 3; ++s) {
                    double va[3], vb[3];
                    for (int r = 0; r < 3; ++r) {
                        va[r] = cbp[s * K + 6 * ba + r];
                        vb[r] = cbp[s * K + 6 * bb + r];
                    }
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) acc[r][c] += va[r] * vb[c];
                }
            }
            for (int s = COMP ? 0 : 3; s < NR; ++s) {
                double va[6], vb[6];
                for (int r = 0; r < 6; ++r) {
                    va[r] = cbp[s * K + 6 * ba + r];
                    vb[r] = cbp[s * K + 6 * bb + r];
                }
                if (split_sh && s >= 6) {
                    if (ba == bb)
                        for (int r = 0; r < 3; ++r)
                            for (int c = 0; c < 3; ++c) accs[r][c] += va[3 + r] * vb[3 + c];
                    continue;
                }
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c) acc[r][c] += va[r] * vb[c];
            }
        }
        if (split_sh) {
            // shear rows of this Gauss point (global basis): factor rows 2q, 2q+1
            for (int i = tid; i < ne * 2 * K; i += WG) {
                const int le = i / (2 * K), rem = i % (2 * K);
                const int rr = rem / K, col = rem % K;
                p.split_cs[(size_t)(e0 + le) * 8 * K + (2 * q + rr) * K + col] = sC[le * CSTR + (6 + rr) * K + col];
            }
        }
    }
    // ---------------- Q4 drilling (FEMMShellQ4RSModule.jl:807-859) ----------------
    if (cact && ba == bb) {
        double *ms = sM + le3 * MISC + 9 * ba;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) ms[r * 3 + c] = acc[3 + r][3 + c] + accs[r][c];
    }
    __syncthreads();
    // trace term of each element node, once per (element, node), by the
    // node's first G3 lane (whose node flag and normal are in registers):
    // projects node k's rotation block onto the plane normal to the nodal
    // normal; -1 marks a node the average skips.  The lane then leaves the
    // node's flag (valid, nonzero normal) and normal in place of the block,
    // for the diagonal-block lanes below.
    if (gact && h2 == 0) {
        const int le = le2, k = j2;
        double *kr = sM + le * MISC + 9 * k;
        double t = -1.0;
        bool counted = false;
        if (valid2) {
            const double nn = sqrt(n20 * n20 + n21 * n21 + n22 * n22);
            if (nn != 0.0) {
                counted = true;
                const double u[3] = {n20 / nn, n21 / nn, n22 / nn};
                double Pm[3][3], tmp[3][3];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) Pm[a][b] = ((a == b) ? 1.0 : 0.0) - u[a] * u[b];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        tmp[a][b] = Pm[a][0] * kr[0 * 3 + b] + Pm[a][1] * kr[1 * 3 + b] + Pm[a][2] * kr[2 * 3 + b];
                double tr = 0.0;
                for (int a = 0; a < 3; ++a) tr += tmp[a][0] * Pm[0][a] + tmp[a][1] * Pm[1][a] + tmp[a][2] * Pm[2][a];
                tr = tr / 2.0;
                t = tr > 0.0 ? tr : 0.0;
            }
        }
        sM[le * MISC + 36 + k] = t;
        kr[0] = counted ? 1.0 : 0.0;
        kr[1] = n20;
        kr[2] = n21;
        kr[3] = n22;
    }
    __syncthreads();
    const int grpd = UG ? 0 : (cact ? (p.elem_group ? p.elem_group[elem(le3)] : 0) : 0);
    const double dscale = COMP ? p.drill_scale : (cact ? p.g_homog[16 * grpd + 15] : 0.0);
    if (split_dr && cact && ba == bb) {
        double *dd = p.split_dr + ((size_t)(e0 + le3) * NEN + ba) * 9;
        for (int i = 0; i < 9; ++i) dd[i] = 0.0;
    }
    if (cact && ba == bb && dscale != 0.0) {
        const int e = elem(le3);
        // kavg: mean over the counted nodes, summed in node order
        double sum = 0.0;
        int cnt = 0;
        for (int k = 0; k < 4; ++k) {
            const double t = sM[le3 * MISC + 36 + k];
            if (t < 0.0) continue;
            sum += t;
            ++cnt;
        }
        const double kavg = (cnt == 0) ? 0.0 : (sum / cnt) * dscale;
        const double *nrec = sM + le3 * MISC + 9 * ba;  // node ba: flag, normal
        if (kavg != 0.0) {
            if (nrec[0] != 0.0) {
                const double nv[3] = {nrec[1], nrec[2], nrec[3]};
                if (split_dr) {
                    double *dd = p.split_dr + ((size_t)e * NEN + ba) * 9;
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) dd[r * 3 + c] = kavg * (nv[r] * nv[c]);
                } else {
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) acc[3 + r][3 + c] += kavg * (nv[r] * nv[c]);
                }
            }
        }
    }
    // ---------------- stage and store ----------------
    if (cact) {
        double *ks = smem + (le3 * NBLK + blk) * 36;  // G and C regions are dead here
        if (ctrans) {
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) ks[c * 6 + r] = acc[r][c];
        } else {
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) ks[r * 6 + c] = acc[r][c];
        }
    }
    __syncthreads();
    if constexpr (FUSED) {
        tile_write<NBLK, WG, 3>(tblk, tp, smem, vals);  // 3 items per pass: measured best of 1-4 on Q4
    } else {
        const int total = ne * NBLK * 36;
        double *dst = Kb + (size_t)e0 * NBLK * 36;
        for (int i = tid; i < total; i += WG) dst[i] = smem[i];
    }
}

// ---------------------------------------------------------------------------
// Scatter: node-pair blocks -> CSR values
// ---------------------------------------------------------------------------
//
// Neighbour slots per (node, row) thread of the slice builder of the
// mixed-precision operator; more neighbours loop.
#define FS_SC_NB 16

// Pair-driven scatter: thread (upper pair (I, J), row component r) sums row r
// of K_IJ over the pair's contributing element blocks in ascending element
// order (a block stored as (J, I) is read transposed) and writes the six values
// into CSR row dof(I, r) at J's column run and, for I != J, the transposed
// entries into rows dof(J, c) at I's column run.  Each element block is read
// once (six threads read its six rows), every CSR value is written exactly
// once, and K is bitwise symmetric by construction.
// 384 threads = 64 node pairs per workgroup (measured against 128, 192, 256
// and 768 on the 512 x 512 cylinder: 0.640, 0.635, 0.628, 0.612 ms; 384: 0.613 ms)
constexpr int SC_WG = 384;
__global__ __launch_bounds__(SC_WG) void scatter_pair_kernel(ScatterPlan sp, const double *__restrict__ Kb,
                                                           double *__restrict__ vals) {
    const int64_t gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int64_t p = sp.pbegin + gid / 6;
    if (p >= sp.pend) return;
    const int r = (int)(gid % 6);
    const int I = sp.pair_node[2 * p], J = sp.pair_node[2 * p + 1];
    const int kI = sp.pair_slot[2 * p], kJ = sp.pair_slot[2 * p + 1];
    double sum[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int ce = sp.contrib_ptr[p + 1];
    for (int q = sp.contrib_ptr[p]; q < ce; ++q) {
        const int cv = sp.contrib[q];
        const double *B = Kb + (int64_t)(cv >> 1) * 36;
        if ((cv & 1) == 0) {
            for (int c = 0; c < 6; ++c) sum[c] += B[r * 6 + c];
        } else {
            for (int c = 0; c < 6; ++c) sum[c] += B[c * 6 + r];
        }
    }
    write_pair_row(sp, I, J, kI, kJ, r, sum, vals);
}

// Boundary pairs of the fused assembly: the six threads of a boundary pair
// add the pair's partial sums in tile order and write its CSR values with
// lane c taking column c of the pair's block for the rows of I and row c for
// the transposed rows of J, so that each store instruction writes six
// consecutive values per pair (a run of J's columns in a row of I, or of I's
// columns in a row of J) instead of six values in six rows.
constexpr int PF_WG = 256;
__global__ __launch_bounds__(PF_WG) void pair_finalize_kernel(ScatterPlan sp, TilePlan tp,
                                                              double *__restrict__ vals) {
    const int64_t gid = xcd_block(blockIdx.x, gridDim.x) * PF_WG + threadIdx.x;
    const int64_t b = gid / 6;
    if (b >= tp.nbnd) return;
    const int c = (int)(gid % 6);
    const int4 ra = tp.brec[3 * b], rb = tp.brec[3 * b + 1], ro = tp.brec[3 * b + 2];
    double col[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, row[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int se = tp.bp_ptr[b + 1];
    for (int k = tp.bp_ptr[b]; k < se; ++k) {
        const double *P = tp.partial + (int64_t)k * 36;
        for (int r = 0; r < 6; ++r) col[r] += P[r * 6 + c];
        for (int j = 0; j < 6; ++j) row[j] += P[c * 6 + j];
    }
    write_pair_lane_rec(ra, rb, ro, c, col, row, vals);
}

// ---------------------------------------------------------------------------
// Nodal normals (associategeometry!)
// ---------------------------------------------------------------------------
//
// Thread per node: the incident elements are visited in ascending order
// (node_elem lists), reproducing the element-order accumulation of
// FEMMShellT3FFModule.jl:579-587 / FEMMShellQ4RSModule.jl:487-495, then the
// normalisation and the validity pass (:597-611 / :508-520).
// Layered form (gnrm set: the constant layup csys normal of each element
// group, FEMMShellT3FFCompModule.jl:500-521 / FEMMShellQ4RSCompModule.jl:
// 462-503): the sum runs over the groups' csys normals (Q4: weighted by the
// corner Jacobian), and validity is checked against the element normal (T3)
// or the csys normal (Q4).
template <int NEN>
__global__ __launch_bounds__(256) void normals_kernel(int64_t N, int E, const double *__restrict__ xs,
                                                      const double *__restrict__ ys,
                                                      const double *__restrict__ zs,
                                                      const int32_t *__restrict__ conn,
                                                      const int32_t *__restrict__ ne_ptr,
                                                      const int32_t *__restrict__ ne_list, double ntol,
                                                      const double *__restrict__ gnrm,
                                                      const int32_t *__restrict__ egroup,
                                                      double *__restrict__ nx, double *__restrict__ ny,
                                                      double *__restrict__ nz, uint8_t *__restrict__ valid) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    double acc[3] = {0.0, 0.0, 0.0};
    const int b = ne_ptr[I], en = ne_ptr[I + 1];
    double nrm[3];
    for (int pass = 0; pass < 2; ++pass) {
        uint8_t ok = 1;
        for (int k = b; k < en; ++k) {
            const int e = ne_list[k];
            double X[NEN][3];
            int loc = 0;
            for (int a = 0; a < NEN; ++a) {
                const int n = conn[a * E + e];
                if (n == I) loc = a;
                X[a][0] = xs[n];
                X[a][1] = ys[n];
                X[a][2] = zs[n];
            }
            double J[3][2], EG[3][3];
            double w = 1.0;
            if (NEN == 3) {
                for (int j = 0; j < 3; ++j) {
                    J[j][0] = X[1][j] - X[0][j];
                    J[j][1] = X[2][j] - X[0][j];
                }
            } else {
                const double npc[4][2] = {{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0}};
                double gNp[4][2];
                q4_gradn(npc[loc][0], npc[loc][1], gNp);
                for (int i = 0; i < 3; ++i)
                    for (int kk = 0; kk < 2; ++kk) {
                        double s = 0.0;
                        for (int a = 0; a < 4; ++a) s += X[a][i] * gNp[a][kk];
                        J[i][kk] = s;
                    }
                const double c0 = -J[2][0] * J[1][1] + J[1][0] * J[2][1];
                const double c1 = J[2][0] * J[0][1] - J[0][0] * J[2][1];
                const double c2 = -J[1][0] * J[0][1] + J[0][0] * J[1][1];
                w = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
            }
            e_g(EG, J);
            // direction summed and checked: the element normal, or the
            // group's csys normal in the layered form
            double dir[3] = {EG[0][2], EG[1][2], EG[2][2]};
            if (gnrm) {
                const double *gn = gnrm + 3 * (egroup ? egroup[e] : 0);
                for (int j = 0; j < 3; ++j) dir[j] = gn[j];
            }
            if (pass == 0) {
                if (NEN == 3) {
                    acc[0] += dir[0];
                    acc[1] += dir[1];
                    acc[2] += dir[2];
                } else {
                    acc[0] += w * dir[0];
                    acc[1] += w * dir[1];
                    acc[2] += w * dir[2];
                }
            } else {
                // checked against the element normal, except layered Q4: the csys normal
                const bool vs_elem = !gnrm || NEN == 3;
                const double nd = vs_elem ? nrm[0] * EG[0][2] + nrm[1] * EG[1][2] + nrm[2] * EG[2][2]
                                          : nrm[0] * dir[0] + nrm[1] * dir[1] + nrm[2] * dir[2];
                if (nd < 1 - ntol) ok = 0;
            }
        }
        if (pass == 0) {
            const double nn = sqrt(acc[0] * acc[0] + acc[1] * acc[1] + acc[2] * acc[2]);
            for (int k = 0; k < 3; ++k) nrm[k] = nn > 0.0 ? acc[k] / nn : acc[k];
        } else {
            nx[I] = nrm[0];
            ny[I] = nrm[1];
            nz[I] = nrm[2];
            valid[I] = ok;
        }
    }
}

// ---------------------------------------------------------------------------
// Lumped mass
// ---------------------------------------------------------------------------
//
// Thread per node: sums the element masses of the incident elements in
// ascending element order.  T3: rho t Ae / 3 and rho t^3/12 Ae / 3
// (FEMMShellT3FFModule.jl:790-791); Q4: 2x2 Gauss sums divided by 4
// (FEMMShellQ4RSModule.jl:994-1002); layered: laminate inertia per group.
struct MassParams {
    int composite;
    const double *g_rho_t;    // homogeneous: ngroups x 2 (rho, thickness)
    const double *g_inertia;  // layered: ngroups x 2
    const int32_t *elem_group;
};

template <int NEN>
__global__ __launch_bounds__(256) void mass_kernel(int64_t N, int E, MassParams mp, const double *__restrict__ xs,
                                                   const double *__restrict__ ys, const double *__restrict__ zs,
                                                   const int32_t *__restrict__ conn,
                                                   const int32_t *__restrict__ ne_ptr,
                                                   const int32_t *__restrict__ ne_list,
                                                   const int64_t *__restrict__ node_dof, int64_t nrows,
                                                   double *__restrict__ diag) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    double tm = 0.0, rm = 0.0;
    bool first = true;
    for (int k = ne_ptr[I]; k < ne_ptr[I + 1]; ++k) {
        const int e = ne_list[k];
        double X[NEN][3];
        for (int a = 0; a < NEN; ++a) {
            const int n = conn[a * E + e];
            X[a][0] = xs[n];
            X[a][1] = ys[n];
            X[a][2] = zs[n];
        }
        double m0 = 0.0, m1 = 0.0, rho = 0.0, tt = 0.0;
        const int g = mp.elem_group ? mp.elem_group[e] : 0;
        if (mp.composite) {
            m0 = mp.g_inertia[2 * g];
            m1 = mp.g_inertia[2 * g + 1];
        } else {
            rho = mp.g_rho_t[2 * g];
            tt = mp.g_rho_t[2 * g + 1];
        }
        double te, re;
        if (NEN == 3) {
            double J0[3][2], EG[3][3];
            for (int j = 0; j < 3; ++j) {
                J0[j][0] = X[1][j] - X[0][j];
                J0[j][1] = X[2][j] - X[0][j];
            }
            e_g(EG, J0);
            const double a = J0[0][0] * EG[0][0] + J0[1][0] * EG[1][0] + J0[2][0] * EG[2][0];
            const double bq = J0[0][0] * EG[0][1] + J0[1][0] * EG[1][1] + J0[2][0] * EG[2][1];
            const double c = J0[0][1] * EG[0][0] + J0[1][1] * EG[1][0] + J0[2][1] * EG[2][0];
            const double d = J0[0][1] * EG[0][1] + J0[1][1] * EG[1][1] + J0[2][1] * EG[2][1];
            const double Ae = (a * d - bq * c) / 2;
            if (!mp.composite) {
                te = rho * (tt * Ae) / 3;
                re = rho * (tt * tt * tt / 12 * Ae) / 3;
            } else {
                te = m0 * (Ae / 3);
                re = m1 * (Ae / 3);
            }
        } else {
            double sa = 0.0, sb = 0.0;
            for (int q = 0; q < 4; ++q) {
                const double xi = (q / 2 == 0) ? -FS_GAUSS2 : FS_GAUSS2;
                const double eta = (q % 2 == 0) ? -FS_GAUSS2 : FS_GAUSS2;
                double gNp[4][2], J[3][2];
                q4_gradn(xi, eta, gNp);
                for (int i = 0; i < 3; ++i)
                    for (int kk = 0; kk < 2; ++kk) {
                        double s = 0.0;
                        for (int a = 0; a < 4; ++a) s += X[a][i] * gNp[a][kk];
                        J[i][kk] = s;
                    }
                const double c0 = -J[2][0] * J[1][1] + J[1][0] * J[2][1];
                const double c1 = J[2][0] * J[0][1] - J[0][0] * J[2][1];
                const double c2 = -J[1][0] * J[0][1] + J[0][0] * J[1][1];
                const double Jac = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
                if (!mp.composite) {
                    sa += rho * tt * Jac * 1.0;
                    sb += rho * (tt * tt * tt) / 12 * Jac * 1.0;
                } else {
                    sa += m0 * Jac * 1.0;
                    sb += m1 * Jac * 1.0;
                }
            }
            te = sa / 4;
            re = sb / 4;
        }
        if (first) {
            tm = te;
            rm = re;
            first = false;
        } else {
            tm += te;
            rm += re;
        }
    }
    for (int c = 0; c < 6; ++c) {
        const int64_t d = node_dof[I * 6 + c];
        if (d >= 0 && d < nrows) diag[d] = c < 3 ? tm : rm;
    }
}

// ---------------------------------------------------------------------------
// Load / boundary-condition / dof front end
// ---------------------------------------------------------------------------
//
// numberdofs! (free dofs first in node-major order, then the fixed ones) as a
// three-pass scan over the N x 6 fixed flags: per-tile free counts, one
// workgroup scanning the tile counts, then each tile renumbers its entries.
// A free entry i gets the number of free entries before it; a fixed entry gets
// nfree + the number of fixed entries before it.
#define FS_DOF_TILE 4096  // entries per workgroup (256 threads x 16)

__global__ __launch_bounds__(256) void dof_count_kernel(int64_t n, const uint8_t *__restrict__ fixed,
                                                        int64_t *__restrict__ tile_free) {
    __shared__ int64_t red[256];
    const int64_t base = (int64_t)blockIdx.x * FS_DOF_TILE + (int64_t)threadIdx.x * 16;
    int64_t cnt = 0;
    for (int k = 0; k < 16; ++k) {
        const int64_t i = base + k;
        if (i < n && !fixed[i]) ++cnt;
    }
    red[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_free[blockIdx.x] = red[0];
}

// Exclusive scan of the tile counts in place by one workgroup; total[0] = nfree.
__global__ __launch_bounds__(256) void dof_scan_kernel(int64_t ntiles, int64_t *__restrict__ tile_free,
                                                       int64_t *__restrict__ total) {
    __shared__ int64_t sh[256];
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < ntiles; b0 += 256) {
        const int64_t i = b0 + threadIdx.x;
        const int64_t v = i < ntiles ? tile_free[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int s = 1; s < 256; s <<= 1) {  // inclusive Hillis-Steele scan
            const int64_t add = (int)threadIdx.x >= s ? sh[threadIdx.x - s] : 0;
            __syncthreads();
            sh[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < ntiles) tile_free[i] = carry + sh[threadIdx.x] - v;
        const int64_t tile_sum = sh[255];
        __syncthreads();
        carry += tile_sum;
    }
    if (threadIdx.x == 0) total[0] = carry;
}

__global__ __launch_bounds__(256) void dof_number_kernel(int64_t n, const uint8_t *__restrict__ fixed,
                                                         const int64_t *__restrict__ tile_free,
                                                         const int64_t *__restrict__ total,
                                                         int64_t *__restrict__ dof) {
    __shared__ int64_t sh[256];
    const int64_t tile0 = (int64_t)blockIdx.x * FS_DOF_TILE;
    const int64_t base = tile0 + (int64_t)threadIdx.x * 16;
    int64_t cnt = 0;
    for (int k = 0; k < 16; ++k) {
        const int64_t i = base + k;
        if (i < n && !fixed[i]) ++cnt;
    }
    sh[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = 1; s < 256; s <<= 1) {
        const int64_t add = (int)threadIdx.x >= s ? sh[threadIdx.x - s] : 0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    int64_t freeb = tile_free[blockIdx.x] + sh[threadIdx.x] - cnt;  // free entries before `base`
    const int64_t nfree = total[0];
    for (int k = 0; k < 16; ++k) {
        const int64_t i = base + k;
        if (i >= n) break;
        if (!fixed[i]) {
            dof[i] = freeb;
            ++freeb;
        } else {
            dof[i] = nfree + (i - freeb);
        }
    }
}

// Integration rules of the surface loads, as FinEtools tabulates them:
// TriRule(1), TriRule(3), TriRule(6) (param_coords, weights) for T3 and
// GaussRule(2, 2) for Q4.
__constant__ double c_tri_rule[10][3] = {
    {1.0 / 3, 1.0 / 3, 1.0 / 2},  // TriRule(1)
    {1.0 / 6, 1.0 / 6, 1.0 / 6},  // TriRule(3)
    {2.0 / 3, 1.0 / 6, 1.0 / 6},
    {1.0 / 6, 2.0 / 3, 1.0 / 6},
    {0.816847572980459, 0.091576213509771, 0.109951743655322 / 2},  // TriRule(6)
    {0.091576213509771, 0.816847572980459, 0.109951743655322 / 2},
    {0.091576213509771, 0.091576213509771, 0.109951743655322 / 2},
    {0.108103018168070, 0.445948490915965, 0.223381589678011 / 2},
    {0.445948490915965, 0.108103018168070, 0.223381589678011 / 2},
    {0.445948490915965, 0.445948490915965, 0.223381589678011 / 2}};

template <int NEN>
__device__ __forceinline__ void load_point(int npts, int q, double N[NEN], double gNp[NEN][2], double &w) {
    if constexpr (NEN == 3) {
        const int r = (npts == 1 ? 0 : npts == 3 ? 1 : 4) + q;
        const double xi = c_tri_rule[r][0], eta = c_tri_rule[r][1];
        N[0] = 1.0 - xi - eta;
        N[1] = xi;
        N[2] = eta;
        gNp[0][0] = -1.0;
        gNp[0][1] = -1.0;
        gNp[1][0] = 1.0;
        gNp[1][1] = 0.0;
        gNp[2][0] = 0.0;
        gNp[2][1] = 1.0;
        w = c_tri_rule[r][2];
    } else {
        const double xi = (q / 2 == 0) ? -FS_GAUSS2 : FS_GAUSS2;
        const double eta = (q % 2 == 0) ? -FS_GAUSS2 : FS_GAUSS2;
        q4_gradn(xi, eta, gNp);
        N[0] = 0.25 * ((1 - xi) * (1 - eta));
        N[1] = 0.25 * ((1 + xi) * (1 - eta));
        N[2] = 0.25 * ((1 + xi) * (1 + eta));
        N[3] = 0.25 * ((1 - xi) * (1 + eta));
        w = 1.0;
    }
}

// Consistent nodal loads of a surface force intensity (distribloads with
// ForceIntensity, FinEtools; e.g. plate_expl_examples.jl:184-203): thread per
// node, summing over its incident elements in ascending order each element's
// entry Fe[a][c] = sum_q (N_a f_c) (Jac w), as the element-order assembly of
// the element vectors does.  force: one 6-vector (per_point 0) or E x npts x 6.
template <int NEN>
__global__ __launch_bounds__(256) void surface_load_kernel(int64_t N, int E, int npts,
                                                           const double *__restrict__ xs,
                                                           const double *__restrict__ ys,
                                                           const double *__restrict__ zs,
                                                           const int32_t *__restrict__ conn,
                                                           const int32_t *__restrict__ ne_ptr,
                                                           const int32_t *__restrict__ ne_list, int per_point,
                                                           const double *__restrict__ force,
                                                           const int64_t *__restrict__ node_dof, int64_t nrows,
                                                           double *__restrict__ F) {
    const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = ne_ptr[I]; k < ne_ptr[I + 1]; ++k) {
        const int e = ne_list[k];
        double X[NEN][3];
        int la = 0;
        for (int a = 0; a < NEN; ++a) {
            const int n = conn[a * E + e];
            if (n == I) la = a;
            X[a][0] = xs[n];
            X[a][1] = ys[n];
            X[a][2] = zs[n];
        }
        double fe[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int q = 0; q < npts; ++q) {
            double Nq[NEN], gNp[NEN][2], w;
            load_point<NEN>(npts, q, Nq, gNp, w);
            double J[3][2];
            for (int i = 0; i < 3; ++i)
                for (int kk = 0; kk < 2; ++kk) {
                    double s = 0.0;
                    for (int a = 0; a < NEN; ++a) s += X[a][i] * gNp[a][kk];
                    J[i][kk] = s;
                }
            const double c0 = -J[2][0] * J[1][1] + J[1][0] * J[2][1];
            const double c1 = J[2][0] * J[0][1] - J[0][0] * J[2][1];
            const double c2 = -J[1][0] * J[0][1] + J[0][0] * J[1][1];
            const double Jw = sqrt(c0 * c0 + c1 * c1 + c2 * c2) * w;
            const double *f = per_point ? force + ((int64_t)e * npts + q) * 6 : force;
            for (int c = 0; c < 6; ++c) fe[c] += (Nq[la] * f[c]) * Jw;
        }
        for (int c = 0; c < 6; ++c) acc[c] += fe[c];
    }
    for (int c = 0; c < 6; ++c) {
        const int64_t d = node_dof[I * 6 + c];
        if (d < nrows) F[d] = acc[c];
    }
}

// Locations of the surface-load integration points, E x npts x 3.
template <int NEN>
__global__ __launch_bounds__(256) void load_points_kernel(int E, int npts, const double *__restrict__ xs,
                                                          const double *__restrict__ ys,
                                                          const double *__restrict__ zs,
                                                          const int32_t *__restrict__ conn,
                                                          double *__restrict__ out) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)E * npts) return;
    const int e = (int)(gid / npts), q = (int)(gid % npts);
    double Nq[NEN], gNp[NEN][2], w;
    load_point<NEN>(npts, q, Nq, gNp, w);
    double p[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < NEN; ++a) {
        const int n = conn[a * E + e];
        p[0] += Nq[a] * xs[n];
        p[1] += Nq[a] * ys[n];
        p[2] += Nq[a] * zs[n];
    }
    for (int k = 0; k < 3; ++k) out[gid * 3 + k] = p[k];
}

// ---------------------------------------------------------------------------
// Stress resultants at the integration points (inspectintegpoints)
// ---------------------------------------------------------------------------
//
// One thread per (listed element, integration point) computes the membrane
// forces, bending moments and transverse shear forces of the point at once
// (FEMMShellT3FFModule.jl:850-962, FEMMShellQ4RSModule.jl:1061-1170,
// FEMMShellT3FFCompModule.jl:809-943, FEMMShellQ4RSCompModule.jl:1077-1209).
// The element displacements are taken into the element frame node by node
// (the block structure of T = Tae Tga: each node's 3 x 3 triad blocks, plus
// the coupling of the in-plane rotations to every node's translations) instead
// of through the dense 18 x 18 / 24 x 24 product, then the strain rows of
// Bm, Bb and Bs are applied, then the material law and the output frame.

// Pieces of t_e = Tae Tga t_g at one point (_TransfmatGToA,
// FEMMShellT3FFModule.jl:398-419; _transfmat_a_to_e!, :421-463), node by
// node: un_j = A_j' E_G' t_j (translations), rn_j = the first two components
// of A_j' E_G' r_j (the drilling component does not enter the rows of T), and
// W0 = sum_j gN_j0 un_j, W1 = sum_j gN_j1 un_j, so that the coupling of node
// i's in-plane rotations to all translations is
// sum_j sum_k a3(i, j, k) un_j[k] = (A_i[1] . W0 - A_i[0] . W1) / 2.
template <int NEN>
__device__ __forceinline__ void nodal_frame(const double EG[3][3], const double A[NEN][3][3],
                                            const double gN[NEN][2], const double tg[NEN][6], double un[NEN][3],
                                            double rn[NEN][2], double W0[3], double W1[3]) {
    for (int k = 0; k < 3; ++k) W0[k] = W1[k] = 0.0;
    for (int j = 0; j < NEN; ++j) {
        double gt[3], gr[3];
        for (int k = 0; k < 3; ++k) {
            gt[k] = EG[0][k] * tg[j][0] + EG[1][k] * tg[j][1] + EG[2][k] * tg[j][2];
            gr[k] = EG[0][k] * tg[j][3] + EG[1][k] * tg[j][4] + EG[2][k] * tg[j][5];
        }
        for (int r = 0; r < 3; ++r) un[j][r] = A[j][0][r] * gt[0] + A[j][1][r] * gt[1] + A[j][2][r] * gt[2];
        for (int r = 0; r < 2; ++r) rn[j][r] = A[j][0][r] * gr[0] + A[j][1][r] * gr[1] + A[j][2][r] * gr[2];
        for (int k = 0; k < 3; ++k) {
            W0[k] += gN[j][0] * un[j][k];
            W1[k] += gN[j][1] * un[j][k];
        }
    }
}

// Element-frame translations te[0..2] and in-plane rotations te[3..4] of node i.
__device__ __forceinline__ void node_element_frame(const double Ai[3][3], const double un[3], const double rn[2],
                                                   const double W0[3], const double W1[3], double te[5]) {
    for (int rw = 0; rw < 3; ++rw) te[rw] = Ai[rw][0] * un[0] + Ai[rw][1] * un[1] + Ai[rw][2] * un[2];
    const double ia33 = 1 / Ai[2][2];
    const double cpl = 0.5 * ((Ai[1][0] * W0[0] + Ai[1][1] * W0[1] + Ai[1][2] * W0[2]) -
                              (Ai[0][0] * W1[0] + Ai[0][1] * W1[1] + Ai[0][2] * W1[2]));
    for (int rw = 0; rw < 2; ++rw) {
        double acc = 0.0;
        for (int cl = 0; cl < 2; ++cl) acc += (Ai[rw][cl] - ia33 * Ai[rw][2] * Ai[cl][2]) * rn[cl];
        te[3 + rw] = acc + (ia33 * Ai[rw][2]) * cpl;
    }
}

// (o2_e' X o2_e) for the symmetric 2 x 2 tensor X = [x0 x2; x2 x1], o2_e = [m n; -n m].
__device__ __forceinline__ void out_tensor(double m, double n, const double x[3], double *o) {
    const double t00 = m * x[0] - n * x[2], t01 = m * x[2] - n * x[1];
    const double t10 = n * x[0] + m * x[2], t11 = n * x[2] + m * x[1];
    o[0] = t00 * m - t01 * n;
    o[1] = t10 * n + t11 * m;
    o[2] = t00 * n + t01 * m;
}

// T3 transverse shear entries bs[node][w, tx, ty][row]: _Bsmat!, the mean of
// the three orderings of _add_Bsmat_o! (FEMMShellT3FFModule.jl:486-537),
// accumulated in the reference's order.
__device__ __forceinline__ void t3_bs(const double ec[3][2], double Ae, double bs[18]) {
    for (int i = 0; i < 18; ++i) bs[i] = 0.0;
    const double mlt = (1 / 2.0 / Ae);
    for (int o = 0; o < 3; ++o) {
        const int s = o, p = (o + 1) % 3, q = (o + 2) % 3;
        const double a = ec[p][0] - ec[s][0], b = ec[p][1] - ec[s][1];
        const double c = ec[q][0] - ec[s][0], d = ec[q][1] - ec[s][1];
        bs[s * 6 + 0 * 2 + 0] += mlt * (b - d);
        bs[s * 6 + 2 * 2 + 0] += mlt * (Ae);
        bs[s * 6 + 0 * 2 + 1] += mlt * (c - a);
        bs[s * 6 + 1 * 2 + 1] += mlt * (-Ae);
        bs[p * 6 + 0 * 2 + 0] += mlt * (d);
        bs[p * 6 + 1 * 2 + 0] += mlt * (-b * d / 2);
        bs[p * 6 + 2 * 2 + 0] += mlt * (a * d / 2);
        bs[p * 6 + 0 * 2 + 1] += mlt * (-c);
        bs[p * 6 + 1 * 2 + 1] += mlt * (b * c / 2);
        bs[p * 6 + 2 * 2 + 1] += mlt * (-a * c / 2);
        bs[q * 6 + 0 * 2 + 0] += mlt * (-b);
        bs[q * 6 + 1 * 2 + 0] += mlt * (b * d / 2);
        bs[q * 6 + 2 * 2 + 0] += mlt * (-b * c / 2);
        bs[q * 6 + 0 * 2 + 1] += mlt * (a);
        bs[q * 6 + 1 * 2 + 1] += mlt * (-a * d / 2);
        bs[q * 6 + 2 * 2 + 1] += mlt * (a * c / 2);
    }
    for (int i = 0; i < 18; ++i) bs[i] *= (1.0 / 3);
}

// TWICE (layered Q4 as the reference computes it): the element-frame vectors
// are transformed by T a second time before the strain rows are applied.
template <int NEN, bool COMP, bool TWICE>
__global__ __launch_bounds__(256) void resultants_kernel(ResParams rp, const double *__restrict__ xs,
                                                         const double *__restrict__ ys,
                                                         const double *__restrict__ zs,
                                                         const int32_t *__restrict__ conn,
                                                         const double *__restrict__ nxs,
                                                         const double *__restrict__ nys,
                                                         const double *__restrict__ nzs,
                                                         const uint8_t *__restrict__ valid) {
    constexpr int NQP = NEN == 3 ? 1 : 4;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)rp.nlist * NQP) return;
    const int64_t l = gid / NQP;
    const int q = (int)(gid % NQP);
    const int E = rp.nelem;
    const int e = rp.felist ? rp.felist[l] : (int)l;
    int nd[NEN];
    for (int a = 0; a < NEN; ++a) nd[a] = conn[a * E + e];
    // ---- geometry of the point: E_G, element-frame gradients, shear entries, h
    double EG[3][3], gN[NEN][2], h;
    double ec[NEN][2], Ae = 0.0;  // element coordinates for the shear rows
    {
        double X[NEN][3];
        for (int a = 0; a < NEN; ++a) {
            X[a][0] = xs[nd[a]];
            X[a][1] = ys[nd[a]];
            X[a][2] = zs[nd[a]];
        }
        if (NEN == 3) {
            double J0[3][2];
            for (int j = 0; j < 3; ++j) {
                J0[j][0] = X[1][j] - X[0][j];
                J0[j][1] = X[2][j] - X[0][j];
            }
            e_g(EG, J0);
            ec[0][0] = ec[0][1] = 0.0;
            for (int k = 0; k < 2; ++k) {
                ec[1][k] = J0[0][0] * EG[0][k] + J0[1][0] * EG[1][k] + J0[2][0] * EG[2][k];
                ec[2][k] = J0[0][1] * EG[0][k] + J0[1][1] * EG[1][k] + J0[2][1] * EG[2][k];
            }
            const double a = ec[1][0], b = ec[1][1], c = ec[2][0], d = ec[2][1];
            const double Jd = a * d - b * c;  // _gradN_e_Ae!, FEMMShellT3FFModule.jl:465-484
            gN[0][0] = (b - d) / Jd;
            gN[1][0] = d / Jd;
            gN[2][0] = -b / Jd;
            gN[0][1] = (c - a) / Jd;
            gN[1][1] = -c / Jd;
            gN[2][1] = a / Jd;
            Ae = Jd / 2;
            h = sqrt(2 * Ae);
        } else {
            const double xi = (q / 2 == 0) ? -FS_GAUSS2 : FS_GAUSS2;
            const double eta = (q % 2 == 0) ? -FS_GAUSS2 : FS_GAUSS2;
            double gNp[4][2];
            q4_gradn(xi, eta, gNp);
            double J[3][2];
            for (int i = 0; i < 3; ++i)
                for (int k = 0; k < 2; ++k) {
                    double acc = 0.0;
                    for (int a = 0; a < 4; ++a) acc += X[a][i] * gNp[a][k];
                    J[i][k] = acc;
                }
            e_g(EG, J);
            double cen[3] = {0.0, 0.0, 0.0};
            for (int a = 0; a < 4; ++a)
                for (int k = 0; k < 3; ++k) cen[k] += X[a][k];
            for (int k = 0; k < 3; ++k) cen[k] /= 4;
            for (int a = 0; a < 4; ++a)
                for (int k = 0; k < 2; ++k) {
                    double dd = 0.0;
                    for (int m = 0; m < 3; ++m) dd += (X[a][m] - cen[m]) * EG[m][k];
                    ec[a][k] = dd;
                }
            const double G00 = J[0][0] * J[0][0] + J[1][0] * J[1][0] + J[2][0] * J[2][0];
            const double G01 = J[0][0] * J[0][1] + J[1][0] * J[1][1] + J[2][0] * J[2][1];
            const double G11 = J[0][1] * J[0][1] + J[1][1] * J[1][1] + J[2][1] * J[2][1];
            const double detG = G00 * G11 - G01 * G01;
            if (detG == 0.0) rp.status[0] = 1;
            const double idetG = 1.0 / detG;
            const double i00 = G11 * idetG, i01 = -G01 * idetG, i11 = G00 * idetG;
            for (int a = 0; a < 4; ++a) {
                const double t0 = i00 * gNp[a][0] + i01 * gNp[a][1];
                const double t1 = i01 * gNp[a][0] + i11 * gNp[a][1];
                double gg[3];
                for (int i = 0; i < 3; ++i) gg[i] = J[i][0] * t0 + J[i][1] * t1;
                gN[a][0] = EG[0][0] * gg[0] + EG[1][0] * gg[1] + EG[2][0] * gg[2];
                gN[a][1] = EG[0][1] * gg[0] + EG[1][1] * gg[1] + EG[2][1] * gg[2];
            }
            double maxd = 0.0;  // _quad_diameter, FEMMShellQ4RSModule.jl:861-870
            for (int i = 1; i < 4; ++i) {
                const double dx = X[i][0] - X[0][0], dy = X[i][1] - X[0][1], dz = X[i][2] - X[0][2];
                const double dist = (dx * dx + dy * dy + dz * dz);
                maxd = dist > maxd ? dist : maxd;
            }
            h = sqrt(maxd);
        }
    }
    // ---- nodal triads, the nodal-frame vectors and the strain rows
    double A[NEN][3][3];
    for (int a = 0; a < NEN; ++a) {
        double nk_e[3];
        if (valid[nd[a]]) {
            const double n0 = nxs[nd[a]], n1 = nys[nd[a]], n2 = nzs[nd[a]];
            for (int i = 0; i < 3; ++i) nk_e[i] = EG[0][i] * n0 + EG[1][i] * n1 + EG[2][i] * n2;
        } else {
            nk_e[0] = 0.0;
            nk_e[1] = 0.0;
            nk_e[2] = 1.0;
        }
        nodal_triad(nk_e, A[a]);
    }
    double un[NEN][3], rn[NEN][2], W0[3], W1[3];
    {
        double tg[NEN][6];
        for (int a = 0; a < NEN; ++a)
            for (int c = 0; c < 6; ++c) tg[a][c] = rp.u[(int64_t)nd[a] * 6 + c];
        nodal_frame<NEN>(EG, A, gN, tg, un, rn, W0, W1);
        if (TWICE) {
            for (int i = 0; i < NEN; ++i) {
                node_element_frame(A[i], un[i], rn[i], W0, W1, tg[i]);
                tg[i][5] = 0.0;
            }
            nodal_frame<NEN>(EG, A, gN, tg, un, rn, W0, W1);
        }
    }
    // strains: Bm, Bb (FEMMShellT3FFModule.jl:539-561), node by node; then
    // Bs from the element coordinates once the triads are dead
    double eps[3] = {0.0, 0.0, 0.0}, kap[3] = {0.0, 0.0, 0.0}, gam[2] = {0.0, 0.0};
    double te[NEN][5];
    for (int i = 0; i < NEN; ++i) {
        node_element_frame(A[i], un[i], rn[i], W0, W1, te[i]);
        eps[0] += gN[i][0] * te[i][0];
        eps[1] += gN[i][1] * te[i][1];
        eps[2] += gN[i][1] * te[i][0] + gN[i][0] * te[i][1];
        kap[0] += gN[i][0] * te[i][4];
        kap[1] += -gN[i][1] * te[i][3];
        kap[2] += -gN[i][0] * te[i][3] + gN[i][1] * te[i][4];
    }
    {
        double bs[6 * NEN];  // [node][w, tx, ty][row]
        if (NEN == 3)
            t3_bs((const double(*)[2])ec, Ae, bs);
        else
            q4_mitc_bs((const double(*)[2])ec, (q / 2 == 0) ? -FS_GAUSS2 : FS_GAUSS2,
                       (q % 2 == 0) ? -FS_GAUSS2 : FS_GAUSS2, bs);
        for (int i = 0; i < NEN; ++i)
            for (int row = 0; row < 2; ++row)
                gam[row] += bs[i * 6 + 0 * 2 + row] * te[i][2] + bs[i * 6 + 1 * 2 + row] * te[i][3] +
                            bs[i * 6 + 2 * 2 + row] * te[i][4];
    }
    // ---- material law
    const int grp = rp.elem_group ? rp.elem_group[e] : 0;
    double f[3], mo[3], qs[2];
    if (!COMP) {
        const double *g = rp.g_mat + 16 * grp;
        const double t = g[13], stab_c = g[14];
        const double stab = (t * t) / (t * t + stab_c * (h * h));
        const double cb = (t * t * t) / 12, cs = t * stab;
        for (int i = 0; i < 3; ++i) {
            f[i] = t * (g[i * 3 + 0] * eps[0] + g[i * 3 + 1] * eps[1] + g[i * 3 + 2] * eps[2]);
            mo[i] = cb * (g[i * 3 + 0] * kap[0] + g[i * 3 + 1] * kap[1] + g[i * 3 + 2] * kap[2]);
        }
        for (int i = 0; i < 2; ++i) qs[i] = cs * (g[9 + i * 2 + 0] * gam[0] + g[9 + i * 2 + 1] * gam[1]);
    } else {
        // layup frame: strains through Tb, resultants back through Tb' (the
        // sA = Tb' A Tb, ... of FEMMShellT3FFCompModule.jl:914-921)
        const double *g = rp.g_mat + 41 * grp;
        const double t = g[40];
        const double stab = (t * t) / (t * t + rp.stab_c * (h * h));
        double m, n;
        layup_angle(EG, g + 31, m, n);
        const double Tb[3][3] = {{m * m, n * n, -(m * n)}, {n * n, m * m, (m * n)},
                                 {2 * (m * n), -2 * (m * n), m * m - n * n}};
        double el[3], kl[3], Nl[3], Ml[3];
        for (int i = 0; i < 3; ++i) {
            el[i] = Tb[i][0] * eps[0] + Tb[i][1] * eps[1] + Tb[i][2] * eps[2];
            kl[i] = Tb[i][0] * kap[0] + Tb[i][1] * kap[1] + Tb[i][2] * kap[2];
        }
        for (int i = 0; i < 3; ++i) {
            double a = 0.0, b = 0.0;
            for (int j = 0; j < 3; ++j) {
                a += g[i * 3 + j] * el[j] + g[9 + i * 3 + j] * kl[j];
                b += g[9 + i * 3 + j] * el[j] + g[18 + i * 3 + j] * kl[j];
            }
            Nl[i] = a;
            Ml[i] = b;
        }
        for (int i = 0; i < 3; ++i) {
            f[i] = Tb[0][i] * Nl[0] + Tb[1][i] * Nl[1] + Tb[2][i] * Nl[2];
            mo[i] = Tb[0][i] * Ml[0] + Tb[1][i] * Ml[1] + Tb[2][i] * Ml[2];
        }
        // sH = Tt' H Tt with Tt = [m -n; n m]
        const double gl0 = m * gam[0] - n * gam[1], gl1 = n * gam[0] + m * gam[1];
        const double Ql0 = g[27] * gl0 + g[28] * gl1, Ql1 = g[29] * gl0 + g[30] * gl1;
        qs[0] = stab * (m * Ql0 + n * Ql1);
        qs[1] = stab * (-n * Ql0 + m * Ql1);
    }
    // ---- output frame
    double L[9];
    if (rp.ocs_kind == 0) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) L[i * 3 + j] = EG[i][j];
    } else {
        const double *src = rp.ocs + (rp.ocs_kind == 1 ? 0 : (l * NQP + q) * 9);
        for (int i = 0; i < 9; ++i) L[i] = src[i];
        if (L[2] * EG[0][2] + L[5] * EG[1][2] + L[8] * EG[2][2] < 0.95) rp.status[1] = 1;
    }
    double om, on;
    layup_angle(EG, L, om, on);
    double *o = rp.out + (l * NQP + q) * 8;
    double ov[8];
    out_tensor(om, on, f, ov);
    out_tensor(om, on, mo, ov + 3);
    ov[6] = om * qs[0] - on * qs[1];
    ov[7] = on * qs[0] + om * qs[1];
    for (int k = 0; k < 8; ++k) o[k] = ov[k];
}

// ---------------------------------------------------------------------------
// Explicit central difference
// ---------------------------------------------------------------------------
//
// The reference step (plate_expl_examples.jl:83-93) is
//   U += dt V + dt^2/2 A;  F = g(t) Fmag;  E = K U;
//   F -= E + C (V + dt/2 A);  V += dt/2 A;  A = invMC F;  V += dt/2 A
// The fused form here keeps the displacement one update ahead: the kernel of
// step n reads U_n (already advanced), finishes the step for its row and
// writes U_{n+1} = U_n + dt V_n + dt^2/2 A_n into the other U buffer, so one
// launch per step reads K once and every state vector once.
//
// cd_init_kernel: A_0 = invMC g_0 Fmag;  U_1 = U_0 + dt V_0 + dt^2/2 A_0.
__global__ __launch_bounds__(256) void cd_init_kernel(int64_t n, double dt, double g0,
                                                      const double *__restrict__ invMC,
                                                      const double *__restrict__ Fmag, const double *__restrict__ U0,
                                                      const double *__restrict__ V, double *__restrict__ A,
                                                      double *__restrict__ U1, const int32_t *__restrict__ row_group,
                                                      const double *__restrict__ group_dt,
                                                      const double *__restrict__ gtab) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double dtr = dt, gr = g0;
    if (row_group) {
        dtr = group_dt[row_group[i]];
        gr = gtab[row_group[i]];
    }
    const double a = invMC[i] * (gr * Fmag[i]);
    A[i] = a;
    const double dt2_2 = ((dtr * dtr) / 2);
    U1[i] = U0[i] + (dtr * V[i] + dt2_2 * a);
}

// One step for a group of TPR lanes per row (TPR in {4, 8, 16, 32, 64}).
// Row R of K times x, lane `lane` of TPR: the lane takes the neighbour runs
// rr.y + lane, rr.y + lane + TPR, ... and sums each run's entries in column
// order; the TPR partial sums are then combined by the caller's butterfly.
// The run records of up to FS_RUN_UNROLL runs are loaded before the values
// and displacements they address.
#ifndef FS_RUN_UNROLL
#define FS_RUN_UNROLL 3
#endif
template <int TPR>
__device__ __forceinline__ double run_row_dot(int4 rr, int lane, const int2 *__restrict__ slotrec,
                                              const double *__restrict__ kval, const double *__restrict__ x) {
    double e = 0.0;
    for (int k0 = rr.y + lane; k0 < rr.z; k0 += TPR * FS_RUN_UNROLL) {
        int2 sr[FS_RUN_UNROLL];
#pragma unroll
        for (int t = 0; t < FS_RUN_UNROLL; ++t) {
            const int k = k0 + t * TPR;
            sr[t] = k < rr.z ? slotrec[k] : make_int2(0, 0);
        }
        double kk[FS_RUN_UNROLL][6], xx[FS_RUN_UNROLL][6];
#pragma unroll
        for (int t = 0; t < FS_RUN_UNROLL; ++t) {
            const int nf = sr[t].x & 15;
            const double *kv = kval + rr.x + (sr[t].x >> 4);
            const double *xv = x + sr[t].y;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                kk[t][c] = c < nf ? kv[c] : 0.0;
                xx[t][c] = c < nf ? xv[c] : 0.0;
            }
        }
#pragma unroll
        for (int t = 0; t < FS_RUN_UNROLL; ++t) {
            const int nf = sr[t].x & 15;
#pragma unroll
            for (int c = 0; c < 6; ++c)
                if (c < nf) e += kk[t][c] * xx[t][c];
        }
    }
    return e;
}

template <int TPR>
__global__ __launch_bounds__(256) void cd_step_kernel(RunRows rows,
                                                      const double *__restrict__ kval, const double *__restrict__ Ucur,
                                                      double *__restrict__ Unext, double *__restrict__ V,
                                                      double *__restrict__ A, const double *__restrict__ invMC,
                                                      const double *__restrict__ Cd, const double *__restrict__ Fmag,
                                                      double gstep, double dt, int damped,
                                                      const int32_t *__restrict__ probe_slot,
                                                      double *__restrict__ probe_dst,
                                                      const int32_t *__restrict__ row_group,
                                                      const double *__restrict__ group_dt,
                                                      const double *__restrict__ gtab) {
    const int64_t gt = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int64_t row = gt / TPR;
    const int lane = (int)(gt % TPR);
    const bool active = row < rows.n;
    double e = 0.0;
    if (active) e = run_row_dot<TPR>(rows.rowrec[row], lane, rows.slotrec, kval, Ucur);
#pragma unroll
    for (int off = TPR / 2; off > 0; off >>= 1) e += __shfl_xor(e, off, TPR);
    if (active && lane == 0) {
        double dtr = dt, gr = gstep;
        if (row_group) {  // batch: the row's model has its own dt and load history
            const int gi = row_group[row];
            dtr = group_dt[gi];
            gr = gtab[gi];
        }
        const double dt_2 = (dtr / 2), dt2_2 = ((dtr * dtr) / 2);
        const double v = V[row], a = A[row];
        double f = gr * Fmag[row];
        if (damped)
            f -= e + Cd[row] * (v + dt_2 * a);
        else
            f -= e;
        double vn = v + dt_2 * a;
        const double an = invMC[row] * f;
        vn += dt_2 * an;
        const double u = Ucur[row];
        V[row] = vn;
        A[row] = an;
        Unext[row] = u + (dtr * vn + dt2_2 * an);
        if (probe_slot) {
            const int ps = probe_slot[row];
            if (ps >= 0) probe_dst[ps] = u;  // peek(step, U, V, t): U of this step
        }
    }
}

// Node-grouped form of the same step: NL lanes per node, for the rows of the
// node's free dofs.  Those rows are consecutive in the free-free block and
// hold the same columns (the runs of the node's neighbours), so they are one
// contiguous stretch of nr x rowlen values.  Lane `lane` takes the positions
// lane, lane + NL, ... of the row: it finds each position's column in the
// run records once, loads that displacement once, and multiplies it with the
// value at that position in each of the nr rows; consecutive lanes read
// consecutive values of a row, so every load instruction reads contiguous
// spans of NL values.  The NL partial sums of each row are combined by a
// butterfly and lane r updates row d0 + r exactly as cd_step_kernel does.
// The sum of a row depends only on positions within the row, so a model
// gives bitwise the same result alone and inside a batch.
// PJ positions per lane are loaded together (the host picks PJ so that NL x
// PJ covers the longest row of the block where it can).
// Row sums of node `rec` (its nr rows, rowlen values each) times x for lane
// `lane` of NL: the lane's positions lane, lane + NL, ..., PJ at a time,
// values of all rows first, then the columns from the run records.
template <int NL, int PJ>
__device__ __forceinline__ void node_rows_dot(int4 rec, int info, int lane, const int2 *__restrict__ slotrec,
                                              const double *__restrict__ kval, const double *__restrict__ Ucur,
                                              double acc[6]) {
    const int64_t kstart = rec.x;
    const int nr = info & 7;
    const int rowlen = info >> 3;
    for (int base = lane; base < rowlen; base += NL * PJ) {
        // values first: their addresses need only the node record
        double kv[6][PJ];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const double *kr = kval + kstart + (int64_t)r * rowlen + base;
#pragma unroll
            for (int j = 0; j < PJ; ++j)
                kv[r][j] = (r < nr && base + j * NL < rowlen) ? kr[j * NL] : 0.0;
        }
        int col[PJ];
#pragma unroll
        for (int j = 0; j < PJ; ++j) col[j] = 0;
        for (int k0 = rec.z; k0 < rec.w; k0 += 4) {
            int2 sr[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) sr[t] = k0 + t < rec.w ? slotrec[k0 + t] : make_int2(0, 0);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int off = sr[t].x >> 4, len = sr[t].x & 15;
#pragma unroll
                for (int j = 0; j < PJ; ++j) {
                    const unsigned d = (unsigned)(base + j * NL - off);
                    if (d < (unsigned)len) col[j] = sr[t].y + (int)d;
                }
            }
        }
        // every position below rowlen lies in exactly one run
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
            const double xv = base + j * NL < rowlen ? Ucur[col[j]] : 0.0;
#pragma unroll
            for (int r = 0; r < 6; ++r) acc[r] += kv[r][j] * xv;
        }
    }
}

// One wavefront per workgroup (measured on the 512 x 512 cylinder: 64, 128,
// 256, 512, 1024 threads gave 158.5, 160.7, 164.0, 184.9, 218.1 us per step).
constexpr int CDN_WG = 64;
template <int NL, int PJ>
__global__ __launch_bounds__(CDN_WG) void cd_node_step_kernel(NodeRows nrw, const int2 *__restrict__ slotrec,
                                                           const double *__restrict__ kval, const double *__restrict__ Ucur,
                                                           double *__restrict__ Unext, double *__restrict__ V,
                                                           double *__restrict__ A, const double *__restrict__ invMC,
                                                           const double *__restrict__ Cd, const double *__restrict__ Fmag,
                                                           double gstep, double dt, int damped,
                                                           const int32_t *__restrict__ probe_slot,
                                                           double *__restrict__ probe_dst,
                                                           const int32_t *__restrict__ row_group,
                                                           const double *__restrict__ group_dt,
                                                           const double *__restrict__ gtab) {
    const int64_t gt = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int64_t node = gt / NL;
    const int lane = (int)(gt % NL);
    const bool active = node < nrw.n;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int nr = 0, d0 = 0;
    if (active) {
        const int4 rec = nrw.rec[node];
        const int info = nrw.info[node];
        d0 = rec.y;
        nr = info & 7;
        node_rows_dot<NL, PJ>(rec, info, lane, slotrec, kval, Ucur, acc);
    }
#pragma unroll
    for (int off = NL / 2; off > 0; off >>= 1)
#pragma unroll
        for (int r = 0; r < 6; ++r) acc[r] += __shfl_xor(acc[r], off, NL);
    if (active && lane < nr) {
        double e = acc[0];
#pragma unroll
        for (int r = 1; r < 6; ++r)
            if (lane == r) e = acc[r];
        const int64_t row = (int64_t)d0 + lane;
        double dtr = dt, gr = gstep;
        if (row_group) {
            const int gi = row_group[row];
            dtr = group_dt[gi];
            gr = gtab[gi];
        }
        const double dt_2 = (dtr / 2), dt2_2 = ((dtr * dtr) / 2);
        const double v = V[row], a = A[row];
        double f = gr * Fmag[row];
        if (damped)
            f -= e + Cd[row] * (v + dt_2 * a);
        else
            f -= e;
        double vn = v + dt_2 * a;
        const double an = invMC[row] * f;
        vn += dt_2 * an;
        const double u = Ucur[row];
        V[row] = vn;
        A[row] = an;
        Unext[row] = u + (dtr * vn + dt2_2 * an);
        if (probe_slot) {
            const int ps = probe_slot[row];
            if (ps >= 0) probe_dst[ps] = u;
        }
    }
}

// ---------------------------------------------------------------------------
// Mixed-precision explicit step (FP32 membrane + bending, FP64 elsewhere)
// ---------------------------------------------------------------------------
//
// K = K_mb + K_s + K_d.  K_mb (membrane, bending and their coupling) is held
// in FP32 in node-block SELL slices (MixedOp); the transverse-shear part
// K_s = sum_e C_e^T C_e stays as FP64 element factor rows C_e (2 rows per T3
// element, 8 per Q4) and the drilling part K_d as an FP64 3 x 3 rotation
// block per node.  The row of dof position R = 6 i + r is
//   (K u)_R = sum_slots K_mb[R, slot] . u_j  +  sum_{e ∋ i} C_e[:, (a, r)]' (C_e u_e)
//             + [r >= 3] D_i[r-3, :] . theta_i
// with the incident elements visited in ascending order, so the result is
// deterministic.  K_mb values are the FP64 element-order sums of the element
// blocks, rounded once to FP32.  In the FP32-shear form (NSR = 0) the slices
// hold K_mb + K_s and the shear term above is absent.

// K_mb slices from the split element blocks: thread per (node, row, neighbour
// slot); the sum over the pair's contributions follows scatter_pair_kernel.
// Split of an FP64 value into FP32 hi and an int16 residual lo in units of
// 2^-15 of hi's ulp: v = hi + lo 2^(e - 165) (e = biased exponent of hi)
// to within 2^(e - 166), a relative 2^-39.  |v - hi| <= ulp(hi) / 2 bounds
// |lo| by 2^14.
__device__ __forceinline__ double mx_scale(float hi) {
    const unsigned e = (__float_as_uint(hi) >> 23) & 0xffu;
    return e ? __longlong_as_double((long long)(e + 858u) << 52) : 0.0;
}

__device__ __forceinline__ void mx_split(double v, float &hi, short &lo) {
    hi = (float)v;
    const double s = mx_scale(hi);
    lo = s != 0.0 ? (short)rint((v - (double)hi) / s) : (short)0;
}

__device__ __forceinline__ double mx_value(float hi, short lo) { return (double)hi + (double)lo * mx_scale(hi); }

__global__ __launch_bounds__(256) void mixed_build_kernel(ScatterPlan sp, const double *__restrict__ Kb,
                                                          const int64_t *__restrict__ goff, float2 *__restrict__ kmb,
                                                          short2 *__restrict__ kml) {
    const int64_t gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int64_t I = gid / (6 * FS_SC_NB);
    if (I >= sp.nnodes) return;
    const int r = (int)((gid / FS_SC_NB) % 6);
    const int kk = (int)(gid % FS_SC_NB);
    float2 *dst = kmb + goff[I / 10] + 6 * (I % 10) + r;
    short2 *dsl = kml + goff[I / 10] + 6 * (I % 10) + r;
    const int nb0 = sp.nbr_ptr[I], nb1 = sp.nbr_ptr[I + 1];
    for (int k = nb0 + kk; k < nb1; k += FS_SC_NB) {
        const int pr = sp.nbr_pair[k];
        const int p = pr >> 1;
        const bool side1 = (pr & 1) != 0;
        double sum[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const int ce = sp.contrib_ptr[p + 1];
        for (int q = sp.contrib_ptr[p]; q < ce; ++q) {
            const int cv = sp.contrib[q];
            const double *B = Kb + (int64_t)(cv >> 1) * 36;
            if (side1 == ((cv & 1) != 0)) {
                for (int c = 0; c < 6; ++c) sum[c] += B[r * 6 + c];
            } else {
                for (int c = 0; c < 6; ++c) sum[c] += B[c * 6 + r];
            }
        }
        const int slot = k - nb0;
        for (int c2 = 0; c2 < 3; ++c2) {
            float h0, h1;
            short l0, l1;
            mx_split(sum[2 * c2], h0, l0);
            mx_split(sum[2 * c2 + 1], h1, l1);
            dst[(3 * slot + c2) * 60] = make_float2(h0, h1);
            dsl[(3 * slot + c2) * 60] = make_short2(l0, l1);
        }
    }
}

// Drilling block per node: sum of the element-node blocks in element order.
template <int NEN>
__global__ __launch_bounds__(256) void mixed_drill_kernel(int64_t N, int E, const int32_t *__restrict__ ne_ptr,
                                                          const int32_t *__restrict__ ne_list,
                                                          const int32_t *__restrict__ conn,
                                                          const double *__restrict__ dre, double *__restrict__ dr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double d[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = ne_ptr[i]; k < ne_ptr[i + 1]; ++k) {
        const int e = ne_list[k];
        int a = 0;
        for (int b = 0; b < NEN; ++b)
            if (conn[b * E + e] == (int32_t)i) a = b;
        const double *src = dre + ((int64_t)e * NEN + a) * 9;
        for (int t = 0; t < 9; ++t) d[t] += src[t];
    }
    double *o = dr + i * 6;
    o[0] = d[0];
    o[1] = d[1];
    o[2] = d[2];
    o[3] = d[4];
    o[4] = d[5];
    o[5] = d[8];
}

// One fused central-difference step on the mixed operator: one wavefront per
// slice of 10 nodes, lane 6 (i mod 10) + r for dof r of node i (lanes 60..63
// idle); the update is the one of cd_step_kernel.
//   K_mb: the row sum runs over the node's neighbours j; each stored value
//     is rebuilt from its FP32 part and int16 residual (a relative 2^-39)
//     and multiplied in FP64.  In the form without factor rows (NSR = 0) the
//     slices carry the transverse shear too.  (A product on displacements
//     relative to node i's rigid motion is exact only where K_mb annihilates
//     rigid rotations, which the nodal-triad transformation of a curved mesh
//     does not; with 39-bit values it is not needed.)
//   K_s (NSR > 0): for every incident element e (ascending), lane r forms the partial
//     sums of s_e = C_e u_e over its own component r; the six lanes of the
//     node add the partials in a fixed order, and each lane adds
//     C_e[:, (a, r)]' s_e.  FP64 throughout.
//   K_d: drilling block times the node's rotations, FP64.
#ifndef MX_CHUNK
#define MX_CHUNK 1  // neighbour slots whose loads are issued together (1: highest occupancy, fastest)
#endif
template <int NEN, int NSR>
__global__ __launch_bounds__(256) void cd_mixed_step_kernel(
    MixedOp op, const double *__restrict__ Ucur, double *__restrict__ Unext, double *__restrict__ V,
    double *__restrict__ A, const double *__restrict__ invMC, const double *__restrict__ Cd,
    const double *__restrict__ Fmag, double gstep, double dt, int damped, int forced,
    const int32_t *__restrict__ probe_slot, double *__restrict__ probe_dst, const int32_t *__restrict__ row_group,
    const double *__restrict__ group_dt, const double *__restrict__ gtab) {
    const int64_t gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int64_t g = gid >> 6;  // slice = wavefront
    const int l = (int)(gid & 63);
    if (l >= 60) return;
    const int64_t i64 = 10 * g + l / 6;
    if (i64 >= op.nnodes) return;
    const int i = (int)i64;
    const int r = l % 6;
    const int64_t R = 6 * i64 + r;
    const int base = (int)(threadIdx.x & 63) - r;  // lane of the node's component 0
    const double2 *ui2 = reinterpret_cast<const double2 *>(Ucur + 6 * i64);  // 16-byte aligned
    const double2 ut2h0 = ui2[1], uh12 = ui2[2];
    const double h0 = ut2h0.y, h1 = uh12.x, h2 = uh12.y;
    const int64_t g0 = op.goff[g];
    const int nsl = (int)((op.goff[g + 1] - g0) / 180);
    const float2 *kv = op.kmb + g0 + l;
    const short2 *kl = op.kml + g0 + l;
    const int nb0 = op.nbr_ptr[i];
    const int cnt = min(op.nbr_ptr[i + 1] - nb0, nsl);
    double e = 0.0;
    // neighbour slots in chunks of MX_CHUNK: all loads of a chunk are issued
    // before its arithmetic
    for (int s0 = 0; s0 < cnt; s0 += MX_CHUNK) {
        int jj[MX_CHUNK];
#pragma unroll
        for (int t = 0; t < MX_CHUNK; ++t) jj[t] = (s0 + t < cnt) ? op.nbr_node[nb0 + s0 + t] : i;
        float2 kk[MX_CHUNK][3];
        short2 ll[MX_CHUNK][3];
        double2 uu[MX_CHUNK][3];
#pragma unroll
        for (int t = 0; t < MX_CHUNK; ++t) {
            const int s = s0 + t;
            if (s < cnt) {
                for (int c2 = 0; c2 < 3; ++c2) {
                    kk[t][c2] = kv[(3 * s + c2) * 60];
                    ll[t][c2] = kl[(3 * s + c2) * 60];
                }
            } else {
                for (int c2 = 0; c2 < 3; ++c2) {
                    kk[t][c2] = make_float2(0.f, 0.f);
                    ll[t][c2] = make_short2(0, 0);
                }
            }
            const double2 *uj = reinterpret_cast<const double2 *>(Ucur + 6 * (int64_t)jj[t]);
            uu[t][0] = uj[0];
            uu[t][1] = uj[1];
            uu[t][2] = uj[2];
        }
#pragma unroll
        for (int t = 0; t < MX_CHUNK; ++t) {
            const double d0 = uu[t][0].x, d1 = uu[t][0].y, d2 = uu[t][1].x;
            const double d3 = uu[t][1].y, d4 = uu[t][2].x, d5 = uu[t][2].y;
            const float2 k0 = kk[t][0], k1 = kk[t][1], k2 = kk[t][2];
            const short2 q0 = ll[t][0], q1 = ll[t][1], q2 = ll[t][2];
            e += mx_value(k0.x, q0.x) * d0 + mx_value(k0.y, q0.y) * d1 + mx_value(k1.x, q1.x) * d2 +
                 mx_value(k1.y, q1.y) * d3 + mx_value(k2.x, q2.x) * d4 + mx_value(k2.y, q2.y) * d5;
        }
    }
    // drilling (FP64)
    if (r >= 3) {
        const double *d = op.dr + i64 * 6;
        const int rr = r - 3;
        const double d0 = rr == 0 ? d[0] : (rr == 1 ? d[1] : d[2]);
        const double d1 = rr == 0 ? d[1] : (rr == 1 ? d[3] : d[4]);
        const double d2 = rr == 0 ? d[2] : (rr == 1 ? d[4] : d[5]);
        e += d0 * h0 + d1 * h1 + d2 * h2;
    }
    if constexpr (NSR > 0) {
        // transverse shear (FP64), incident elements in ascending order; the six
        // lanes of the node share the element products through cross-lane reads.
        // Incidence records: element id and its nodes (ne_pack, stride 4 or 8).
        constexpr int PS = NEN == 3 ? 4 : 8;
        constexpr int K = 6 * NEN;
        for (int k = op.ne_ptr[i]; k < op.ne_ptr[i + 1]; ++k) {
            const int4 pk = reinterpret_cast<const int4 *>(op.ne_pack + (int64_t)k * PS)[0];
            int nodes[NEN];
            nodes[0] = pk.y;
            nodes[1] = pk.z;
            nodes[2] = pk.w;
            if constexpr (NEN == 4) nodes[3] = op.ne_pack[(int64_t)k * PS + 4];
            const double *C = op.cs + (int64_t)pk.x * NSR * K;
            double ub[NEN];
            for (int b2 = 0; b2 < NEN; ++b2) ub[b2] = Ucur[6 * (int64_t)nodes[b2] + r];
            double part[NSR], ca[NSR];
            for (int q = 0; q < NSR; ++q) {
                part[q] = 0.0;
                ca[q] = 0.0;
            }
            for (int b2 = 0; b2 < NEN; ++b2) {
                for (int q = 0; q < NSR; ++q) {
                    const double cqb = C[q * K + 6 * b2 + r];
                    part[q] += cqb * ub[b2];
                    if (nodes[b2] == i) ca[q] = cqb;  // column (a, r) of the node's own block
                }
            }
            for (int q = 0; q < NSR; ++q) {
                double sq = 0.0;
                for (int c = 0; c < 6; ++c) sq += __shfl(part[q], base + c, 64);
                e += ca[q] * sq;
            }
        }
    }
    double dtr = dt, gr = gstep;
    if (row_group) {
        const int gi = row_group[R];
        dtr = group_dt[gi];
        gr = gtab[gi];
    }
    const double dt_2 = (dtr / 2), dt2_2 = ((dtr * dtr) / 2);
    const double v = V[R], a0 = A[R];
    double f = forced ? gr * Fmag[R] : 0.0;
    if (damped)
        f -= e + Cd[R] * (v + dt_2 * a0);
    else
        f -= e;
    double vn = v + dt_2 * a0;
    const double an = invMC[R] * f;
    vn += dt_2 * an;
    const double u = Ucur[R];
    V[R] = vn;
    A[R] = an;
    Unext[R] = u + (dtr * vn + dt2_2 * an);
    if (probe_slot) {
        const int ps = probe_slot[R];
        if (ps >= 0) probe_dst[ps] = u;
    }
}

// probes: out[k] = U[probe[k]]
__global__ void gather_kernel(int n, const int64_t *__restrict__ idx, const double *__restrict__ U,
                              double *__restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = U[idx[k]];
}

// y = K x (power iteration)
template <int TPR>
__global__ __launch_bounds__(256) void spmv_kernel(RunRows rows, const double *__restrict__ kval,
                                                   const double *__restrict__ x, double *__restrict__ y) {
    const int64_t gt = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int64_t row = gt / TPR;
    const int lane = (int)(gt % TPR);
    double e = 0.0;
    if (row < rows.n) e = run_row_dot<TPR>(rows.rowrec[row], lane, rows.slotrec, kval, x);
#pragma unroll
    for (int off = TPR / 2; off > 0; off >>= 1) e += __shfl_xor(e, off, TPR);
    if (row < rows.n && lane == 0) y[row] = e;
}

// y = K x in the node-grouped form of cd_node_step_kernel (same row sums).
template <int NL, int PJ>
__global__ __launch_bounds__(CDN_WG) void spmv_node_kernel(NodeRows nrw, const int2 *__restrict__ slotrec,
                                                        const double *__restrict__ kval, const double *__restrict__ x,
                                                        double *__restrict__ y) {
    const int64_t gt = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int64_t node = gt / NL;
    const int lane = (int)(gt % NL);
    const bool active = node < nrw.n;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int nr = 0, d0 = 0;
    if (active) {
        const int4 rec = nrw.rec[node];
        const int info = nrw.info[node];
        d0 = rec.y;
        nr = info & 7;
        node_rows_dot<NL, PJ>(rec, info, lane, slotrec, kval, x, acc);
    }
#pragma unroll
    for (int off = NL / 2; off > 0; off >>= 1)
#pragma unroll
        for (int r = 0; r < 6; ++r) acc[r] += __shfl_xor(acc[r], off, NL);
    if (active && lane < nr) {
        double e = acc[0];
#pragma unroll
        for (int r = 1; r < 6; ++r)
            if (lane == r) e = acc[r];
        y[(int64_t)d0 + lane] = e;
    }
}

// Deterministic reductions for the power iteration: fixed tree per block,
// then one block sums the block partials in index order.
// mode 0: sum x*x; mode 1: sum x*y; mode 2: sum x*(d*x)
__global__ __launch_bounds__(256) void dot_partial_kernel(int64_t n, int mode, const double *__restrict__ x,
                                                          const double *__restrict__ y,
                                                          const double *__restrict__ d, double *__restrict__ part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double xv = x[i];
        s += mode == 0 ? xv * xv : (mode == 1 ? xv * y[i] : xv * (d[i] * xv));
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void dot_final_kernel(int nb, const double *__restrict__ part,
                                                        double *__restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// Start of the power iteration: inv[i] = 1 / mdiag[i] (the invM of pwr,
// clamp_cyl_transient_examples.jl:47) and the four scale factors set to one
// (a fill kernel rather than a memset, so that the sequence can be captured).
__global__ __launch_bounds__(256) void pwr_init_kernel(int64_t n, const double *__restrict__ mdiag,
                                                       double *__restrict__ inv, double *__restrict__ ss) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4) ss[i] = 1.0;
    if (i < n) inv[i] = 1.0 / mdiag[i];
}

// y = (d ? d : 1) * x / sqrt(ss[0])  (normalisation steps of pwr)
__global__ __launch_bounds__(256) void scale_kernel(int64_t n, const double *__restrict__ ss,
                                                    const double *__restrict__ d, const double *__restrict__ x,
                                                    double *__restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double f = 1.0 / sqrt(ss[0]);
    y[i] = d ? d[i] * (x[i] * f) : x[i] * f;
}

// ---------------------------------------------------------------------------
// Launch wrappers (host side of this translation unit)
// ---------------------------------------------------------------------------

// lanes per node in the G2/G3 phases of the T3 kernel (1 or 2)
#ifndef FS_T3_RS
#define FS_T3_RS 2
#endif

// ---------------------------------------------------------------------------
// Generic element-matrix reduction (assemble! / makematrix! of a FEMM with no
// element kernel here)
// ---------------------------------------------------------------------------
//
// Thread per stored entry of the CSR pattern: the entry's contributions are
// positions in the buffered element matrices, listed in the order in which
// assemble! appended them (element order, column-major within an element),
// and are summed in that order, as sparse(I, J, V) combines duplicates.
__global__ __launch_bounds__(256) void coo_reduce_kernel(int64_t nnz, const int64_t *__restrict__ cptr,
                                                        const int64_t *__restrict__ clist,
                                                        const double *__restrict__ mats, double *__restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    const int64_t q0 = cptr[s], q1 = cptr[s + 1];
    double acc = mats[clist[q0]];
    for (int64_t q = q0 + 1; q < q1; ++q) acc += mats[clist[q]];
    out[s] = acc;
}

template <int NEN, int NQP, int NR, int NE, int WG, int SPLIT, int RS, bool FUSED, bool COMP>
static hipError_t launch_elem_kind(const ElemParams &p, const double *x, const double *y, const double *z,
                              const int32_t *conn, const double *nx, const double *ny, const double *nz,
                              const uint8_t *valid, double *Kb, int *status, hipStream_t st,
                              const ScatterPlan *sp = nullptr, const TilePlan *tp = nullptr, double *vals = nullptr) {
    using S = ElemSmem<NEN, NQP, NR, NE>;
    const size_t smem = S::TOTAL * sizeof(double);
    static_assert(S::TOTAL * sizeof(double) <= 160 * 1024, "element kernel LDS exceeds 160 KiB");
    static_assert(NE * NQP * NEN * RS <= WG && NE * (NEN * (NEN + 1) / 2) <= WG, "workgroup too small");
    // set on every launch: cheap, per device, and free of shared host state
    // when several contexts launch from different threads
    if (smem > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)elem_kernel<NEN, NQP, NR, NE, WG, SPLIT, RS, FUSED, COMP>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        if (e != hipSuccess) return e;
    }
    int nblocks;
    if constexpr (FUSED) {
        if (!sp || !tp || !vals || !p.tile_elem || !p.tile_conn || p.ebegin != 0) return hipErrorInvalidValue;
        nblocks = (int)tp->ntiles;
    } else {
        if (p.ebegin < 0 || p.eend > p.nelem || p.ebegin > p.eend) return hipErrorInvalidValue;
        nblocks = (p.eend - p.ebegin + NE - 1) / NE;
    }
    if (nblocks == 0) return hipSuccess;
    hipLaunchKernelGGL((elem_kernel<NEN, NQP, NR, NE, WG, SPLIT, RS, FUSED, COMP>), dim3(nblocks), dim3(WG), smem, st, p, x,
                       y, z, conn, nx, ny, nz, valid, Kb, status, sp ? *sp : ScatterPlan{}, tp ? *tp : TilePlan{},
                       vals);
    return hipGetLastError();
}

// The T3 kernel is compiled once per material kind (homogeneous / layered),
// so that the membrane-row skip and the kind's branches resolve at compile
// time and the accumulation over the rows of C unrolls.
template <int NEN, int NQP, int NR, int NE, int WG, int SPLIT, int RS = FS_T3_RS, bool FUSED = false>
static hipError_t launch_elem(const ElemParams &p, const double *x, const double *y, const double *z,
                              const int32_t *conn, const double *nx, const double *ny, const double *nz,
                              const uint8_t *valid, double *Kb, int *status, hipStream_t st,
                              const ScatterPlan *sp = nullptr, const TilePlan *tp = nullptr, double *vals = nullptr) {
    if (p.composite)
        return launch_elem_kind<NEN, NQP, NR, NE, WG, SPLIT, RS, FUSED, true>(p, x, y, z, conn, nx, ny, nz, valid, Kb,
                                                                             status, st, sp, tp, vals);
    return launch_elem_kind<NEN, NQP, NR, NE, WG, SPLIT, RS, FUSED, false>(p, x, y, z, conn, nx, ny, nz, valid, Kb,
                                                                          status, st, sp, tp, vals);
}

// lanes per node in the G2/G3 phases of the Q4 kernel (1 or 2)
#ifndef FS_Q4_RS
#define FS_Q4_RS 2
#endif

template <int NE, int WG, int SPLIT, int RS, bool COMP, bool FUSED, bool UG>
static hipError_t launch_q4_seq_ug(const ElemParams &p, const double *x, const double *y, const double *z,
                                const int32_t *conn, const double *nx, const double *ny, const double *nz,
                                const uint8_t *valid, double *Kb, int *status, hipStream_t st,
                                const ScatterPlan *sp = nullptr, const TilePlan *tp = nullptr, double *vals = nullptr) {
    // geometry records (homogeneous 44, layered 47 doubles per Gauss point),
    // one Gauss point of C, node records (also the drilling scratch)
    constexpr int total = NE * 4 * (COMP ? 47 : 44) + NE * (8 * 24 + 2) + NE * 4 * 14;
    const size_t smem = total * sizeof(double);
    static_assert(NE * 10 <= WG && NE * 4 * RS <= WG, "workgroup too small");
    static_assert(total * sizeof(double) <= 160 * 1024, "element kernel LDS exceeds 160 KiB");
    if (smem > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)elem_q4_seq_kernel<NE, WG, SPLIT, RS, COMP, FUSED, UG>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        if (e != hipSuccess) return e;
    }
    int nblocks;
    if constexpr (FUSED) {
        if (!sp || !tp || !vals || !p.tile_elem || !p.tile_conn || p.ebegin != 0) return hipErrorInvalidValue;
        nblocks = (int)tp->ntiles;
    } else {
        if (p.ebegin < 0 || p.eend > p.nelem || p.ebegin > p.eend) return hipErrorInvalidValue;
        nblocks = (p.eend - p.ebegin + NE - 1) / NE;
    }
    if (nblocks == 0) return hipSuccess;
    hipLaunchKernelGGL((elem_q4_seq_kernel<NE, WG, SPLIT, RS, COMP, FUSED, UG>), dim3(nblocks), dim3(WG), smem, st, p, x, y,
                       z, conn, nx, ny, nz, valid, Kb, status, sp ? *sp : ScatterPlan{}, tp ? *tp : TilePlan{}, vals);
    return hipGetLastError();
}

// Models with one material group (no per-element group array) run the
// uniform-group instantiation; the split forms (built once per stiffness
// for the mixed step) keep the general one.
template <int NE, int WG, int SPLIT, int RS, bool COMP, bool FUSED = false>
static hipError_t launch_q4_seq(const ElemParams &p, const double *x, const double *y, const double *z,
                                const int32_t *conn, const double *nx, const double *ny, const double *nz,
                                const uint8_t *valid, double *Kb, int *status, hipStream_t st,
                                const ScatterPlan *sp = nullptr, const TilePlan *tp = nullptr, double *vals = nullptr) {
    if (SPLIT == 0 && p.elem_group == nullptr)
        return launch_q4_seq_ug<NE, WG, SPLIT, RS, COMP, FUSED, SPLIT == 0>(p, x, y, z, conn, nx, ny, nz, valid, Kb,
                                                                             status, st, sp, tp, vals);
    return launch_q4_seq_ug<NE, WG, SPLIT, RS, COMP, FUSED, false>(p, x, y, z, conn, nx, ny, nz, valid, Kb, status,
                                                                   st, sp, tp, vals);
}

// Launch shapes (elements per workgroup, workgroup size) the autotuner
// chooses from; every shape computes bitwise the same blocks.
int element_variants(int nen) { return nen == 3 ? 3 : 3; }

void element_variant_shape(int nen, int variant, int *ne, int *wg) {
    static const int t3[3][2] = {{16, 128}, {8, 64}, {32, 256}};
    static const int q4[3][2] = {{16, 192}, {12, 128}, {6, 64}};
    const int v = (variant >= 0 && variant < 3) ? variant : 0;
    *ne = nen == 3 ? t3[v][0] : q4[v][0];
    *wg = nen == 3 ? t3[v][1] : q4[v][1];
}

hipError_t launch_element_blocks(const ElemParams &p, const double *x, const double *y, const double *z,
                                 const int32_t *conn, const double *nx, const double *ny, const double *nz,
                                 const uint8_t *valid, double *Kb, int *status, hipStream_t st, int variant) {
#define FS_ARGS p, x, y, z, conn, nx, ny, nz, valid, Kb, status, st
    // split modes of the mixed-precision operator: 1 shear factor rows and
    // drilling blocks out, 2 drilling blocks out only
    if (p.split_dr != nullptr) {
        if (p.split_cs != nullptr) {
            if (p.nen == 3) {
                if (p.formulation == 1) return launch_elem<3, 1, 12, 16, 128, 1>(FS_ARGS);
                return launch_elem<3, 1, 8, 16, 128, 1>(FS_ARGS);
            }
            return p.composite ? launch_q4_seq<16, 192, 1, FS_Q4_RS, true>(FS_ARGS)
                               : launch_q4_seq<16, 192, 1, FS_Q4_RS, false>(FS_ARGS);
        }
        if (p.nen == 3) {
            if (p.formulation == 1) return launch_elem<3, 1, 12, 16, 128, 2>(FS_ARGS);
            return launch_elem<3, 1, 8, 16, 128, 2>(FS_ARGS);
        }
        return p.composite ? launch_q4_seq<16, 192, 2, FS_Q4_RS, true>(FS_ARGS)
                           : launch_q4_seq<16, 192, 2, FS_Q4_RS, false>(FS_ARGS);
    }
    if (p.nen == 3) {
        if (p.formulation == 1) {
            switch (variant) {
                case 1: return launch_elem<3, 1, 12, 8, 64, 0>(FS_ARGS);
                case 2: return launch_elem<3, 1, 12, 32, 256, 0>(FS_ARGS);
                default: return launch_elem<3, 1, 12, 16, 128, 0>(FS_ARGS);
            }
        }
        switch (variant) {
            case 1: return launch_elem<3, 1, 8, 8, 64, 0>(FS_ARGS);
            case 2: return launch_elem<3, 1, 8, 32, 256, 0>(FS_ARGS);
            default: return launch_elem<3, 1, 8, 16, 128, 0>(FS_ARGS);
        }
    }
    switch (variant) {
        case 1: return p.composite ? launch_q4_seq<12, 128, 0, FS_Q4_RS, true>(FS_ARGS)
                                   : launch_q4_seq<12, 128, 0, FS_Q4_RS, false>(FS_ARGS);
        case 2: return p.composite ? launch_q4_seq<6, 64, 0, FS_Q4_RS, true>(FS_ARGS)
                                   : launch_q4_seq<6, 64, 0, FS_Q4_RS, false>(FS_ARGS);
        default: return p.composite ? launch_q4_seq<16, 192, 0, FS_Q4_RS, true>(FS_ARGS)
                                    : launch_q4_seq<16, 192, 0, FS_Q4_RS, false>(FS_ARGS);
    }
#undef FS_ARGS
}

// Tile size of the fused assembly per family (elements per workgroup of the
// fused element kernels).
int fused_tile_elems(int nen) { return nen == 3 ? 32 : 12; }

hipError_t launch_element_fused(const ElemParams &p, const double *x, const double *y, const double *z,
                                const int32_t *conn, const double *nx, const double *ny, const double *nz,
                                const uint8_t *valid, int *status, const ScatterPlan &sp, const TilePlan &tp,
                                double *vals, hipStream_t st) {
#define FS_FARGS p, x, y, z, conn, nx, ny, nz, valid, nullptr, status, st, &sp, &tp, vals
    if (p.split_dr != nullptr || p.split_cs != nullptr) return hipErrorInvalidValue;
    if (p.nen == 3) {
        if (p.formulation == 1) return launch_elem<3, 1, 12, 32, 256, 0, FS_T3_RS, true>(FS_FARGS);
        return launch_elem<3, 1, 8, 32, 256, 0, FS_T3_RS, true>(FS_FARGS);
    }
    return p.composite ? launch_q4_seq<12, 128, 0, FS_Q4_RS, true, true>(FS_FARGS)
                       : launch_q4_seq<12, 128, 0, FS_Q4_RS, false, true>(FS_FARGS);
#undef FS_FARGS
}

hipError_t launch_pair_finalize(const ScatterPlan &sp, const TilePlan &tp, double *vals, hipStream_t st) {
    const int64_t threads = tp.nbnd * 6;
    if (threads == 0) return hipSuccess;
    hipLaunchKernelGGL(pair_finalize_kernel, dim3((unsigned)((threads + PF_WG - 1) / PF_WG)), dim3(PF_WG), 0, st, sp,
                       tp, vals);
    return hipGetLastError();
}

hipError_t launch_scatter(const ScatterPlan &sp, const double *Kb, double *vals, hipStream_t st) {
    if (sp.pbegin < 0 || sp.pend > sp.npairs || sp.pbegin > sp.pend) return hipErrorInvalidValue;
    const int64_t threads = (sp.pend - sp.pbegin) * 6;
    if (threads == 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_pair_kernel, dim3((unsigned)((threads + SC_WG - 1) / SC_WG)), dim3(SC_WG), 0, st, sp, Kb,
                       vals);
    return hipGetLastError();
}

hipError_t launch_normals(int nen, int64_t N, int E, const double *x, const double *y, const double *z,
                          const int32_t *conn, const int32_t *ne_ptr, const int32_t *ne_list, double ntol,
                          const double *gnrm, const int32_t *egroup, double *nx, double *ny, double *nz,
                          uint8_t *valid, hipStream_t st) {
    if (N == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((N + 255) / 256);
    if (nen == 3)
        hipLaunchKernelGGL(normals_kernel<3>, dim3(blocks), dim3(256), 0, st, N, E, x, y, z, conn, ne_ptr, ne_list,
                           ntol, gnrm, egroup, nx, ny, nz, valid);
    else
        hipLaunchKernelGGL(normals_kernel<4>, dim3(blocks), dim3(256), 0, st, N, E, x, y, z, conn, ne_ptr, ne_list,
                           ntol, gnrm, egroup, nx, ny, nz, valid);
    return hipGetLastError();
}

hipError_t launch_number_dofs(int64_t n, const uint8_t *fixed, int64_t *tile_free, int64_t *total, int64_t *dof,
                              hipStream_t st) {
    if (n == 0) return hipSuccess;
    const int64_t ntiles = (n + FS_DOF_TILE - 1) / FS_DOF_TILE;
    hipLaunchKernelGGL(dof_count_kernel, dim3((unsigned)ntiles), dim3(256), 0, st, n, fixed, tile_free);
    hipLaunchKernelGGL(dof_scan_kernel, dim3(1), dim3(256), 0, st, ntiles, tile_free, total);
    hipLaunchKernelGGL(dof_number_kernel, dim3((unsigned)ntiles), dim3(256), 0, st, n, fixed, tile_free, total, dof);
    return hipGetLastError();
}

int64_t number_dofs_tiles(int64_t n) { return (n + FS_DOF_TILE - 1) / FS_DOF_TILE; }

hipError_t launch_surface_load(int nen, int npts, int64_t N, int E, const double *x, const double *y,
                               const double *z, const int32_t *conn, const int32_t *ne_ptr, const int32_t *ne_list,
                               int per_point, const double *force, const int64_t *node_dof, int64_t nrows, double *F,
                               hipStream_t st) {
    if (N == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((N + 255) / 256);
    if (nen == 3)
        hipLaunchKernelGGL(surface_load_kernel<3>, dim3(blocks), dim3(256), 0, st, N, E, npts, x, y, z, conn, ne_ptr,
                           ne_list, per_point, force, node_dof, nrows, F);
    else
        hipLaunchKernelGGL(surface_load_kernel<4>, dim3(blocks), dim3(256), 0, st, N, E, npts, x, y, z, conn, ne_ptr,
                           ne_list, per_point, force, node_dof, nrows, F);
    return hipGetLastError();
}

hipError_t launch_load_points(int nen, int npts, int E, const double *x, const double *y, const double *z,
                              const int32_t *conn, double *out, hipStream_t st) {
    const int64_t nt = (int64_t)E * npts;
    if (nt == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((nt + 255) / 256);
    if (nen == 3)
        hipLaunchKernelGGL(load_points_kernel<3>, dim3(blocks), dim3(256), 0, st, E, npts, x, y, z, conn, out);
    else
        hipLaunchKernelGGL(load_points_kernel<4>, dim3(blocks), dim3(256), 0, st, E, npts, x, y, z, conn, out);
    return hipGetLastError();
}

hipError_t launch_resultants(int nen, int composite, const ResParams &rp, const double *x, const double *y,
                             const double *z, const int32_t *conn, const double *nx, const double *ny,
                             const double *nz, const uint8_t *valid, hipStream_t st) {
    const int64_t nthreads = (int64_t)rp.nlist * (nen == 3 ? 1 : 4);
    if (nthreads == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((nthreads + 255) / 256);
#define FS_RES_CASE(NN, CC, TW)                                                                               \
    if (nen == NN && (composite != 0) == CC && (rp.twice != 0) == TW) {                                         \
        hipLaunchKernelGGL((resultants_kernel<NN, CC, TW>), dim3(blocks), dim3(256), 0, st, rp, x, y, z, conn, nx, \
                           ny, nz, valid);                                                                      \
        return hipGetLastError();                                                                               \
    }
    FS_RES_CASE(3, false, false)
    FS_RES_CASE(3, true, false)
    FS_RES_CASE(4, false, false)
    FS_RES_CASE(4, true, false)
    FS_RES_CASE(4, true, true)
#undef FS_RES_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_mass(int nen, int64_t N, int E, int composite, const double *g_rho_t, const double *g_inertia,
                       const int32_t *elem_group, const double *x, const double *y, const double *z,
                       const int32_t *conn, const int32_t *ne_ptr, const int32_t *ne_list, const int64_t *node_dof,
                       int64_t nrows, double *diag, hipStream_t st) {
    if (N == 0) return hipSuccess;
    MassParams mp{composite, g_rho_t, g_inertia, elem_group};
    const unsigned blocks = (unsigned)((N + 255) / 256);
    if (nen == 3)
        hipLaunchKernelGGL(mass_kernel<3>, dim3(blocks), dim3(256), 0, st, N, E, mp, x, y, z, conn, ne_ptr, ne_list,
                           node_dof, nrows, diag);
    else
        hipLaunchKernelGGL(mass_kernel<4>, dim3(blocks), dim3(256), 0, st, N, E, mp, x, y, z, conn, ne_ptr, ne_list,
                           node_dof, nrows, diag);
    return hipGetLastError();
}

hipError_t launch_cd_init(int64_t n, double dt, double g0, const double *invMC, const double *Fmag, const double *U0,
                          const double *V, double *A, double *U1, const int32_t *row_group, const double *group_dt,
                          const double *gtab, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(cd_init_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, dt, g0, invMC, Fmag,
                       U0, V, A, U1, row_group, group_dt, gtab);
    return hipGetLastError();
}

hipError_t launch_cd_step(int tpr, const RunRows &rows, const double *kval,
                          const double *Ucur, double *Unext, double *V, double *A, const double *invMC,
                          const double *Cd, const double *Fmag, double gstep, double dt, int damped,
                          const int32_t *probe_slot, double *probe_dst, const int32_t *row_group,
                          const double *group_dt, const double *gtab, hipStream_t st) {
    if (rows.n == 0) return hipSuccess;
    const int64_t threads = rows.n * tpr;
    const unsigned blocks = (unsigned)((threads + 255) / 256);
#define FS_CD_CASE(T)                                                                                           \
    case T:                                                                                                     \
        hipLaunchKernelGGL(cd_step_kernel<T>, dim3(blocks), dim3(256), 0, st, rows, kval, Ucur,                  \
                           Unext, V, A, invMC, Cd, Fmag, gstep, dt, damped, probe_slot, probe_dst, row_group,    \
                           group_dt, gtab);                                                                    \
        break;
    switch (tpr) {
        FS_CD_CASE(4)
        FS_CD_CASE(8)
        FS_CD_CASE(16)
        FS_CD_CASE(32)
        FS_CD_CASE(64)
        default:
            return hipErrorInvalidValue;
    }
#undef FS_CD_CASE
    return hipGetLastError();
}

hipError_t launch_cd_node_step(int nl, int pj, const NodeRows &nrw, const int2 *slotrec, const double *kval,
                               const double *Ucur, double *Unext, double *V, double *A, const double *invMC,
                               const double *Cd, const double *Fmag, double gstep, double dt, int damped,
                               const int32_t *probe_slot, double *probe_dst, const int32_t *row_group,
                               const double *group_dt, const double *gtab, hipStream_t st) {
    if (nrw.n == 0) return hipSuccess;
    const int64_t threads = nrw.n * nl;
    const unsigned blocks = (unsigned)((threads + CDN_WG - 1) / CDN_WG);
#define FS_CDN_CASE(T, P)                                                                                       \
    if (nl == T && pj == P) {                                                                                   \
        hipLaunchKernelGGL((cd_node_step_kernel<T, P>), dim3(blocks), dim3(CDN_WG), 0, st, nrw, slotrec, kval, Ucur, \
                           Unext, V, A, invMC, Cd, Fmag, gstep, dt, damped, probe_slot, probe_dst, row_group,    \
                           group_dt, gtab);                                                                    \
        return hipGetLastError();                                                                               \
    }
    FS_CDN_CASE(8, 4)
    FS_CDN_CASE(8, 5)
    FS_CDN_CASE(8, 6)
    FS_CDN_CASE(8, 7)
    FS_CDN_CASE(8, 8)
    FS_CDN_CASE(16, 2)
    FS_CDN_CASE(16, 3)
    FS_CDN_CASE(16, 4)
#undef FS_CDN_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_gather(int n, const int64_t *idx, const double *U, double *out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, idx, U, out);
    return hipGetLastError();
}

hipError_t launch_spmv(const RunRows &rows, const double *kval, const double *x, double *y, hipStream_t st) {
    if (rows.n == 0) return hipSuccess;
    const int64_t threads = rows.n * 4;
    hipLaunchKernelGGL(spmv_kernel<4>, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, rows, kval, x, y);
    return hipGetLastError();
}

hipError_t launch_spmv_node(int nl, int pj, const NodeRows &nrw, const int2 *slotrec, const double *kval,
                            const double *x, double *y, hipStream_t st) {
    if (nrw.n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((nrw.n * nl + CDN_WG - 1) / CDN_WG);
#define FS_SPN_CASE(T, P)                                                                                         \
    if (nl == T && pj == P) {                                                                                     \
        hipLaunchKernelGGL((spmv_node_kernel<T, P>), dim3(blocks), dim3(CDN_WG), 0, st, nrw, slotrec, kval, x, y);   \
        return hipGetLastError();                                                                                 \
    }
    FS_SPN_CASE(8, 4)
    FS_SPN_CASE(8, 5)
    FS_SPN_CASE(8, 6)
    FS_SPN_CASE(8, 7)
    FS_SPN_CASE(8, 8)
    FS_SPN_CASE(16, 2)
    FS_SPN_CASE(16, 3)
    FS_SPN_CASE(16, 4)
#undef FS_SPN_CASE
    return hipErrorInvalidValue;
}

static constexpr int kDotBlocks = 512;

hipError_t launch_dot(int64_t n, int mode, const double *x, const double *y, const double *d, double *part,
                      double *out, hipStream_t st) {
    hipLaunchKernelGGL(dot_partial_kernel, dim3(kDotBlocks), dim3(256), 0, st, n, mode, x, y, d, part);
    hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(256), 0, st, kDotBlocks, part, out);
    return hipGetLastError();
}

hipError_t launch_pwr_init(int64_t n, const double *mdiag, double *inv, double *ss, hipStream_t st) {
    const int64_t m = n > 4 ? n : 4;
    hipLaunchKernelGGL(pwr_init_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, n, mdiag, inv, ss);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Workspace of the allocation-free entry points
// ---------------------------------------------------------------------------
// One copy per device, held until the process ends (NOALLOC_DOUBLES FP64).
// Every call on a device shares it: calls that overlap on one device race.
// Host code must not use the array's name as a device address: it reaches
// the device copy through the address hipGetSymbolAddress returns with that
// device current, taken once per device by noalloc_workspace.
__device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[256] = {};
}  // namespace

hipError_t noalloc_workspace(int device, double **out) {
    *out = nullptr;
    if (device < 0 || device >= 256) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

hipError_t launch_scale(int64_t n, const double *ss, const double *d, const double *x, double *y, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, ss, d, x, y);
    return hipGetLastError();
}

hipError_t launch_mixed_build(const ScatterPlan &sp, const double *Kb, const int64_t *goff, float2 *kmb,
                              short2 *kml, hipStream_t st) {
    const int64_t threads = sp.nnodes * 6 * FS_SC_NB;
    if (threads == 0) return hipSuccess;
    hipLaunchKernelGGL(mixed_build_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, sp, Kb, goff,
                       kmb, kml);
    return hipGetLastError();
}

hipError_t launch_mixed_drill(int nen, int64_t N, int E, const int32_t *ne_ptr, const int32_t *ne_list,
                              const int32_t *conn, const double *dre, double *dr, hipStream_t st) {
    if (N == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((N + 255) / 256);
    if (nen == 3)
        hipLaunchKernelGGL(mixed_drill_kernel<3>, dim3(blocks), dim3(256), 0, st, N, E, ne_ptr, ne_list, conn, dre, dr);
    else
        hipLaunchKernelGGL(mixed_drill_kernel<4>, dim3(blocks), dim3(256), 0, st, N, E, ne_ptr, ne_list, conn, dre, dr);
    return hipGetLastError();
}

hipError_t launch_coo_reduce(int64_t nnz, const int64_t *cptr, const int64_t *clist, const double *mats,
                             double *out, hipStream_t st) {
    if (nnz == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((nnz + 255) / 256);
    hipLaunchKernelGGL(coo_reduce_kernel, dim3(blocks), dim3(256), 0, st, nnz, cptr, clist, mats, out);
    return hipGetLastError();
}

hipError_t launch_cd_mixed_step(int nen, int nsr, const MixedOp &op, const double *Ucur, double *Unext,
                                double *V, double *A, const double *invMC, const double *Cd, const double *Fmag,
                                double gstep, double dt, int damped, int forced, const int32_t *probe_slot,
                                double *probe_dst, const int32_t *row_group, const double *group_dt,
                                const double *gtab, hipStream_t st) {
    if (op.nnodes == 0) return hipSuccess;
    const int64_t threads = (int64_t)((op.nnodes + 9) / 10) * 64;  // one wavefront per 10-node slice
    const unsigned blocks = (unsigned)((threads + 255) / 256);
#define FS_MX_CASE(NEN_, NSR_)                                                                                  \
    if ((nen == NEN_ || NSR_ == 0) && nsr == NSR_) {                                                            \
        hipLaunchKernelGGL((cd_mixed_step_kernel<NEN_, NSR_>), dim3(blocks), dim3(256), 0, st, op, Ucur,         \
                           Unext, V, A, invMC, Cd, Fmag, gstep, dt, damped, forced, probe_slot, probe_dst,       \
                           row_group, group_dt, gtab);                                                          \
        return hipGetLastError();                                                                               \
    }
    FS_MX_CASE(3, 0)  // FP32 shear in the slices (either family)
    FS_MX_CASE(3, 2)
    FS_MX_CASE(3, 6)
    FS_MX_CASE(4, 8)
#undef FS_MX_CASE
    return hipErrorInvalidValue;
}

}  // namespace fs
