/*
 * shell_oracle.c -- CPU oracle for the shell finite-element hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference algorithms (FinEtoolsFlexStructures.jl v3.6.4 plus the FinEtools /
 * FinEtoolsDeforLinear primitives it calls).  It is the checker that the HIP
 * product path is compared against; it is never linked into, loaded by or
 * called from the product library.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py use it.
 *
 * The restatement deliberately follows the reference's own formulation (dense
 * 18x18 / 24x24 element matrices, dense Q^T E Q transforms, add_btdb_ut_only!
 * + complete_lt!, COO append followed by duplicate summation in element order)
 * so that it is an independent computation of the same numbers the product
 * computes in a structure-aware way on the GPU.
 *
 * Conventions: all matrices are row-major C arrays, 0-based.  A reference
 * index (i, j) (1-based, Julia column-major) is element [i-1][j-1] here.
 *
 * Pinning (what checks each function): see the header of oracle/oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NDOF 6

/* ------------------------------------------------------------------------ */
/* FinEtools MatrixUtilityModule / RotationUtilModule primitives (unvendored, */
/* restated from their documented semantics and the call sites).              */
/* ------------------------------------------------------------------------ */

/* add_btdb_ut_only!(Ke, B, Jac_w, D, DB): Ke[upper] += B' * (Jac_w * D * B).
 * Call sites: FEMMShellT3FFModule.jl:681,683,689,699. */
static void add_btdb_ut_only(int kd, double *Ke, int nstr, const double *B,
                             double jac_w, const double *D) {
    double DB[12 * 24];
    for (int nx = 0; nx < kd; ++nx)
        for (int px = 0; px < nstr; ++px) {
            double acc = 0.0;
            for (int mx = 0; mx < nstr; ++mx) acc += D[px * nstr + mx] * B[mx * kd + nx];
            DB[px * kd + nx] = jac_w * acc;
        }
    for (int nx = 0; nx < kd; ++nx)
        for (int mx = 0; mx <= nx; ++mx) {
            double acc = 0.0;
            for (int px = 0; px < nstr; ++px) acc += B[px * kd + mx] * DB[px * kd + nx];
            Ke[mx * kd + nx] += acc;
        }
}

/* add_b1tdb2!(Ke, B1, B2, Jac_w, D, DB): Ke += B1' * (Jac_w * D * B2), all
 * entries.  Call sites: FEMMShellT3FFCompModule.jl:633-634. */
static void add_b1tdb2(int kd, double *Ke, int nstr, const double *B1, const double *B2,
                       double jac_w, const double *D) {
    double DB[12 * 24];
    for (int nx = 0; nx < kd; ++nx)
        for (int px = 0; px < nstr; ++px) {
            double acc = 0.0;
            for (int mx = 0; mx < nstr; ++mx) acc += D[px * nstr + mx] * B2[mx * kd + nx];
            DB[px * kd + nx] = jac_w * acc;
        }
    for (int nx = 0; nx < kd; ++nx)
        for (int mx = 0; mx < kd; ++mx) {
            double acc = 0.0;
            for (int px = 0; px < nstr; ++px) acc += B1[px * kd + mx] * DB[px * kd + nx];
            Ke[mx * kd + nx] += acc;
        }
}

/* complete_lt!: copy the upper triangle into the lower triangle. */
static void complete_lt(int kd, double *Ke) {
    for (int nx = 0; nx < kd; ++nx)
        for (int mx = nx + 1; mx < kd; ++mx) Ke[mx * kd + nx] = Ke[nx * kd + mx];
}

/* TransformerQtEQ (TransformerModule.jl:35-42): E <- Q' * (E * Q). */
static void qteq(int kd, double *E, const double *Q) {
    double buf[24 * 24];
    for (int i = 0; i < kd; ++i)
        for (int j = 0; j < kd; ++j) {
            double acc = 0.0;
            for (int k = 0; k < kd; ++k) acc += E[i * kd + k] * Q[k * kd + j];
            buf[i * kd + j] = acc;
        }
    for (int i = 0; i < kd; ++i)
        for (int j = 0; j < kd; ++j) {
            double acc = 0.0;
            for (int k = 0; k < kd; ++k) acc += Q[k * kd + i] * buf[k * kd + j];
            E[i * kd + j] = acc;
        }
}

/* small dense product C = A * B for n x n blocks */
static void matmul(int n, const double *A, const double *B, double *C) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc += A[i * n + k] * B[k * n + j];
            C[i * n + j] = acc;
        }
}

static double norm3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* rotmat3! (FinEtools RotationUtilModule): Rodrigues rotation matrix for the
 * rotation vector a: R = cos|a| (I - aa') + sin|a|/|a| skew(a) + aa'. */
static void rotmat3(double R[3][3], const double a[3]) {
    double na = norm3(a);
    double S[3][3] = {{0.0, -a[2], a[1]}, {a[2], 0.0, -a[0]}, {-a[1], a[0], 0.0}};
    double u[3] = {a[0] / na, a[1] / na, a[2] / na};
    double ca = cos(na), sa = sin(na);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double aa = u[i] * u[j];
            double id = (i == j) ? 1.0 : 0.0;
            R[i][j] = ca * (id - aa) + (sa / na) * S[i][j] + aa;
        }
}

/* ------------------------------------------------------------------------ */
/* Shared shell kinematics                                                    */
/* ------------------------------------------------------------------------ */

/* _e_g! (FEMMShellT3FFModule.jl:283-305; FEMMShellQ4RSModule.jl:248-264). */
static void e_g(double EG[3][3], const double J[3][2]) {
    for (int j = 0; j < 3; ++j) EG[j][0] = J[j][0];
    double n = sqrt(EG[0][0] * EG[0][0] + EG[1][0] * EG[1][0] + EG[2][0] * EG[2][0]);
    for (int j = 0; j < 3; ++j) EG[j][0] /= n;
    EG[0][2] = -EG[2][0] * J[1][1] + EG[1][0] * J[2][1];
    EG[1][2] = EG[2][0] * J[0][1] - EG[0][0] * J[2][1];
    EG[2][2] = -EG[1][0] * J[0][1] + EG[0][0] * J[1][1];
    n = sqrt(EG[0][2] * EG[0][2] + EG[1][2] * EG[1][2] + EG[2][2] * EG[2][2]);
    for (int j = 0; j < 3; ++j) EG[j][2] /= n;
    EG[0][1] = -EG[2][2] * EG[1][0] + EG[1][2] * EG[2][0];
    EG[1][1] = EG[2][2] * EG[0][0] - EG[0][2] * EG[2][0];
    EG[2][1] = -EG[1][2] * EG[0][0] + EG[0][2] * EG[1][0];
}

/* _NodalTriadsE (FEMMShellT3FFModule.jl:355-388): nodal triads on the element
 * basis.  nrm: nn x 3 nodal normals of the element's nodes; valid: nn flags. */
static void nodal_triads_e(int nn, double AEs[][3][3], int *nvalid, const double EG[3][3],
                           const double *nrm, const uint8_t *valid) {
    for (int k = 0; k < nn; ++k) {
        double nk_e[3];
        nvalid[k] = valid[k] ? 1 : 0;
        if (nvalid[k]) {
            for (int i = 0; i < 3; ++i)
                nk_e[i] = EG[0][i] * nrm[k * 3 + 0] + EG[1][i] * nrm[k * 3 + 1] + EG[2][i] * nrm[k * 3 + 2];
        } else {
            nk_e[0] = 0.0; nk_e[1] = 0.0; nk_e[2] = 1.0;
        }
        /* cross3!(r, f3_e, nk_e) with f3_e = (0, 0, 1) */
        double r[3];
        r[0] = -1.0 * nk_e[1] + 0.0 * nk_e[2];
        r[1] = 1.0 * nk_e[0] - 0.0 * nk_e[2];
        r[2] = -0.0 * nk_e[0] + 0.0 * nk_e[1];
        if (norm3(r) > 1.0e-12) {
            rotmat3(AEs[k], r);
        } else {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) AEs[k][i][j] = (i == j) ? 1.0 : 0.0;
        }
    }
}

/* _TransfmatGToA (FEMMShellT3FFModule.jl:398-419): block diagonal A_E' E_G'. */
static void transfmat_g_to_a(int nn, double *T, double AEs[][3][3], const double EG[3][3]) {
    int kd = nn * NDOF;
    memset(T, 0, sizeof(double) * kd * kd);
    for (int i = 0; i < nn; ++i) {
        double blk[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double acc = 0.0;
                for (int k = 0; k < 3; ++k) acc += AEs[i][k][r] * EG[c][k];
                blk[r][c] = acc;
            }
        int off = i * NDOF;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                T[(off + r) * kd + off + c] = blk[r][c];
                T[(off + 3 + r) * kd + off + 3 + c] = blk[r][c];
            }
    }
}

/* _transfmat_a_to_e! (FEMMShellT3FFModule.jl:421-463). gradN: nn x 2. */
static void transfmat_a_to_e(int nn, double *T, double AEs[][3][3], const double *gradN) {
    int kd = nn * NDOF;
    memset(T, 0, sizeof(double) * kd * kd);
    for (int i = 0; i < nn; ++i) {
        int ro = i * NDOF;
        double (*A)[3] = AEs[i];
        double a33 = A[2][2];
        for (int cl = 0; cl < 3; ++cl)
            for (int rw = 0; rw < 3; ++rw) T[(ro + rw) * kd + ro + cl] = A[rw][cl];
        for (int cl = 0; cl < 2; ++cl)
            for (int rw = 0; rw < 2; ++rw)
                T[(ro + 3 + rw) * kd + ro + 3 + cl] = A[rw][cl] - (1.0 / a33) * A[rw][2] * A[cl][2];
        double m1 = (1.0 / a33) * A[0][2];
        double m2 = (1.0 / a33) * A[1][2];
        for (int j = 0; j < nn; ++j) {
            int co = j * NDOF;
            for (int k = 0; k < 3; ++k) {
                double a3 = 1.0 / 2.0 * (A[1][k] * gradN[j * 2 + 0] - A[0][k] * gradN[j * 2 + 1]);
                T[(ro + 3) * kd + co + k] += m1 * a3;
                T[(ro + 4) * kd + co + k] += m2 * a3;
            }
        }
    }
}

/* Layup2ElementAngle (TransformerModule.jl:92-105). */
static void layup2element_angle(const double EG[3][3], const double L[3][3], double *m, double *n) {
    double M[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += EG[k][i] * L[k][j];
            M[i][j] = acc;
        }
    double M2[2][2] = {{M[0][0], M[0][1]}, {M[1][0], M[1][1]}};
    double c0 = sqrt(M2[0][0] * M2[0][0] + M2[1][0] * M2[1][0]);
    M2[0][0] /= c0; M2[1][0] /= c0;
    double c1 = sqrt(M2[0][1] * M2[0][1] + M2[1][1] * M2[1][1]);
    M2[0][1] /= c1; M2[1][1] /= c1;
    double mm = (M2[0][0] + M2[1][1]) / 2;
    double nn = (M2[0][1] - M2[1][0]) / 2;
    double sn = (nn >= 0.0) ? 1.0 : -1.0;
    *m = mm;
    *n = sn * sqrt(1 - mm * mm);
}

/* plane_stress_Tbar_matrix! (CompositeLayupModule.jl:364-378) and the
 * transformations used per element (FEMMShellT3FFCompModule.jl:619-626). */
static void rotate_laminate(double m, double n, const double *A, const double *B, const double *D,
                            const double *H, double *sA, double *sB, double *sD, double *sH) {
    double Tinv[9] = {m * m, n * n, 2 * (m * n), n * n, m * m, -2 * (m * n), -(m * n), (m * n), m * m - n * n};
    double Tb[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Tb[i * 3 + j] = Tinv[j * 3 + i];
    memcpy(sA, A, 9 * sizeof(double)); qteq(3, sA, Tb);
    memcpy(sB, B, 9 * sizeof(double)); qteq(3, sB, Tb);
    memcpy(sD, D, 9 * sizeof(double)); qteq(3, sD, Tb);
    double Tt[4] = {m, -n, n, m};
    memcpy(sH, H, 4 * sizeof(double)); qteq(2, sH, Tt);
}

/* ------------------------------------------------------------------------ */
/* T3FF / T3FFComp element stiffness                                          */
/* ------------------------------------------------------------------------ */

/* _add_Bsmat_o! (FEMMShellT3FFModule.jl:486-526) -- Bs is 2 x 18. */
static void add_bsmat_o(double *Bs, const double ec[3][2], double Ae, int s, int p, int q) {
    double a = ec[p][0] - ec[s][0];
    double b = ec[p][1] - ec[s][1];
    double c = ec[q][0] - ec[s][0];
    double d = ec[q][1] - ec[s][1];
    double m = (1.0 / 2 / Ae);
    int co = s * 6;
    Bs[0 * 18 + co + 2] += m * (b - d);
    Bs[0 * 18 + co + 4] += m * (Ae);
    Bs[1 * 18 + co + 2] += m * (c - a);
    Bs[1 * 18 + co + 3] += m * (-Ae);
    co = p * 6;
    Bs[0 * 18 + co + 2] += m * (d);
    Bs[0 * 18 + co + 3] += m * (-b * d / 2);
    Bs[0 * 18 + co + 4] += m * (a * d / 2);
    Bs[1 * 18 + co + 2] += m * (-c);
    Bs[1 * 18 + co + 3] += m * (b * c / 2);
    Bs[1 * 18 + co + 4] += m * (-a * c / 2);
    co = q * 6;
    Bs[0 * 18 + co + 2] += m * (-b);
    Bs[0 * 18 + co + 3] += m * (b * d / 2);
    Bs[0 * 18 + co + 4] += m * (-b * c / 2);
    Bs[1 * 18 + co + 2] += m * (a);
    Bs[1 * 18 + co + 3] += m * (-a * d / 2);
    Bs[1 * 18 + co + 4] += m * (a * c / 2);
}

/* Geometry of a T3 facet: FEMMShellT3FFModule.jl:269-321,465-484. */
static void t3_geometry(const double *X, double J0[3][2], double EG[3][3], double ec[3][2],
                        double gradN[3][2], double *Ae, double centroid[3]) {
    for (int k = 0; k < 3; ++k) centroid[k] = (X[0 * 3 + k] + X[1 * 3 + k] + X[2 * 3 + k]) / 3;
    for (int j = 0; j < 3; ++j) {
        J0[j][0] = X[1 * 3 + j] - X[0 * 3 + j];
        J0[j][1] = X[2 * 3 + j] - X[0 * 3 + j];
    }
    e_g(EG, J0);
    ec[0][0] = ec[0][1] = 0.0;
    ec[1][0] = J0[0][0] * EG[0][0] + J0[1][0] * EG[1][0] + J0[2][0] * EG[2][0];
    ec[1][1] = J0[0][0] * EG[0][1] + J0[1][0] * EG[1][1] + J0[2][0] * EG[2][1];
    ec[2][0] = J0[0][1] * EG[0][0] + J0[1][1] * EG[1][0] + J0[2][1] * EG[2][0];
    ec[2][1] = J0[0][1] * EG[0][1] + J0[1][1] * EG[1][1] + J0[2][1] * EG[2][1];
    double a = ec[1][0] - ec[0][0];
    double b = ec[1][1] - ec[0][1];
    double c = ec[2][0] - ec[0][0];
    double d = ec[2][1] - ec[0][1];
    double J = (a * d - b * c);
    gradN[0][0] = (b - d) / J;
    gradN[1][0] = d / J;
    gradN[2][0] = -b / J;
    gradN[0][1] = (c - a) / J;
    gradN[1][1] = -c / J;
    gradN[2][1] = a / J;
    *Ae = J / 2;
}

/*
 * T3FF element stiffness (FEMMShellT3FFModule.jl:670-733) and T3FFComp
 * element stiffness (FEMMShellT3FFCompModule.jl:606-685).
 *   X      3 x 3 node coordinates (row = node)
 *   nrm    3 x 3 nodal normals of the element's nodes; valid: 3 flags
 *   composite == 0: mat = [Dps(9), Dt(4)] (Dt WITHOUT the 5/6 factor; the
 *                   factor is applied here as at :658-659), t = thickness
 *   composite == 1: mat = [A(9), B(9), D(9), H(4)] of the layup, t = layup
 *                   thickness, csys = 3 x 3 layup csys matrix at the centroid
 *   stab_c: stabilization t^2 / (t^2 + stab_c h^2)
 *   formulation: 0 = AVERAGE_B, 1 = AVERAGE_K
 *   Ke (out): 18 x 18 row-major, global basis.
 */
void or_t3ff_elmat(const double *X, const double *nrm, const uint8_t *valid, int composite,
                   const double *mat, const double *csys, double t, double stab_c, int formulation,
                   double drill_scale, double *Ke) {
    const int kd = 18;
    double J0[3][2], EG[3][3], ec[3][2], gradN[3][2], Ae, centroid[3];
    t3_geometry(X, J0, EG, ec, gradN, &Ae, centroid);
    double Dps[9], Dt[4], sA[9], sB[9], sD[9], sH[4];
    if (!composite) {
        memcpy(Dps, mat, 9 * sizeof(double));
        memcpy(Dt, mat + 9, 4 * sizeof(double));
        double scf = 5.0 / 6.0;
        for (int i = 0; i < 4; ++i) Dt[i] *= scf;
    } else {
        double m, n;
        layup2element_angle(EG, (const double(*)[3])csys, &m, &n);
        rotate_laminate(m, n, mat, mat + 9, mat + 18, mat + 27, sA, sB, sD, sH);
    }
    memset(Ke, 0, sizeof(double) * kd * kd);
    double Bm[3 * 18], Bb[3 * 18], Bs[2 * 18];
    memset(Bm, 0, sizeof(Bm));
    memset(Bb, 0, sizeof(Bb));
    for (int i = 0; i < 3; ++i) {
        /* _Bmmat! :539-548 */
        Bm[0 * 18 + 6 * i + 0] = gradN[i][0];
        Bm[1 * 18 + 6 * i + 1] = gradN[i][1];
        Bm[2 * 18 + 6 * i + 0] = gradN[i][1];
        Bm[2 * 18 + 6 * i + 1] = gradN[i][0];
        /* _Bbmat! :550-561 */
        Bb[0 * 18 + 6 * i + 4] = gradN[i][0];
        Bb[1 * 18 + 6 * i + 3] = -gradN[i][1];
        Bb[2 * 18 + 6 * i + 3] = -gradN[i][0];
        Bb[2 * 18 + 6 * i + 4] = gradN[i][1];
    }
    double h = sqrt(2 * Ae);
    double stab = (t * t) / (t * t + stab_c * (h * h));
    if (!composite) {
        add_btdb_ut_only(kd, Ke, 3, Bm, t * Ae, Dps);
        add_btdb_ut_only(kd, Ke, 3, Bb, (t * t * t) / 12 * Ae, Dps);
    } else {
        add_btdb_ut_only(kd, Ke, 3, Bm, Ae, sA);
        add_btdb_ut_only(kd, Ke, 3, Bb, Ae, sD);
        add_b1tdb2(kd, Ke, 3, Bm, Bb, Ae, sB);
        add_b1tdb2(kd, Ke, 3, Bb, Bm, Ae, sB);
    }
    const double *Ds = composite ? sH : Dt;
    double wfac = composite ? stab : t * stab;
    if (formulation == 1) {
        int ord[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
        for (int o = 0; o < 3; ++o) {
            memset(Bs, 0, sizeof(Bs));
            add_bsmat_o(Bs, ec, Ae, ord[o][0], ord[o][1], ord[o][2]);
            add_btdb_ut_only(kd, Ke, 2, Bs, wfac * Ae / 3, Ds);
        }
    } else {
        memset(Bs, 0, sizeof(Bs));
        add_bsmat_o(Bs, ec, Ae, 0, 1, 2);
        add_bsmat_o(Bs, ec, Ae, 1, 2, 0);
        add_bsmat_o(Bs, ec, Ae, 2, 0, 1);
        for (int i = 0; i < 36; ++i) Bs[i] *= (1.0 / 3);
        add_btdb_ut_only(kd, Ke, 2, Bs, wfac * Ae, Ds);
    }
    complete_lt(kd, Ke);
    double AEs[3][3][3];
    int nvalid[3];
    double T[18 * 18];
    nodal_triads_e(3, AEs, nvalid, EG, nrm, valid);
    transfmat_a_to_e(3, T, AEs, &gradN[0][0]);
    qteq(kd, Ke, T);
    double kavg = (((((Ke[3 * kd + 3] + Ke[9 * kd + 9]) + Ke[15 * kd + 15]) + Ke[4 * kd + 4]) +
                    Ke[10 * kd + 10]) + Ke[16 * kd + 16]) / 6 * drill_scale;
    if (nvalid[0]) Ke[5 * kd + 5] += kavg;
    if (nvalid[1]) Ke[11 * kd + 11] += kavg;
    if (nvalid[2]) Ke[17 * kd + 17] += kavg;
    transfmat_g_to_a(3, T, AEs, EG);
    qteq(kd, Ke, T);
}

/* ------------------------------------------------------------------------ */
/* Q4RS / Q4RSComp element stiffness                                          */
/* ------------------------------------------------------------------------ */

/* GaussRule(2, 2) as FinEtools lays it out: 15-digit abscissa, i outer, j inner. */
static const double GAUSS2 = 0.577350269189626;

static void q4_bfun(double xi, double eta, double N[4], double gN[4][2]) {
    N[0] = 0.25 * (1. - xi) * (1. - eta);
    N[1] = 0.25 * (1. + xi) * (1. - eta);
    N[2] = 0.25 * (1. + xi) * (1. + eta);
    N[3] = 0.25 * (1. - xi) * (1. + eta);
    gN[0][0] = -(1. - eta) * 0.25; gN[0][1] = -(1. - xi) * 0.25;
    gN[1][0] = (1. - eta) * 0.25;  gN[1][1] = -(1. + xi) * 0.25;
    gN[2][0] = (1. + eta) * 0.25;  gN[2][1] = (1. + xi) * 0.25;
    gN[3][0] = -(1. + eta) * 0.25; gN[3][1] = (1. - xi) * 0.25;
}

static void q4_gauss_point(int j, double *xi, double *eta) {
    double pc[2] = {-GAUSS2, GAUSS2};
    *xi = pc[j / 2];
    *eta = pc[j % 2];
}

/* locjac! (FinEtools): J = ecoords' * gradNparams (3 x 2). */
static void q4_jac(const double *X, double gN[4][2], double J[3][2]) {
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 2; ++k) {
            double acc = 0.0;
            for (int a = 0; a < 4; ++a) acc += X[a * 3 + i] * gN[a][k];
            J[i][k] = acc;
        }
}

/* Jacobiansurface for a 2-manifold in 3D: || J[:,1] x J[:,2] || */
static double surf_jac(double J[3][2]) {
    double c[3];
    c[0] = -J[2][0] * J[1][1] + J[1][0] * J[2][1];
    c[1] = J[2][0] * J[0][1] - J[0][0] * J[2][1];
    c[2] = -J[1][0] * J[0][1] + J[0][0] * J[1][1];
    return norm3(c);
}

/* _Bsmat MITC (FEMMShellQ4RSModule.jl:753-804) -- Bs_loc 2 x 24 before T. */
static void q4_bsmat_loc(double *Bs, const double ec[4][2], double r, double s) {
    memset(Bs, 0, sizeof(double) * 48);
    double X1 = ec[0][0], Y1 = ec[0][1], X2 = ec[1][0], Y2 = ec[1][1];
    double X3 = ec[2][0], Y3 = ec[2][1], X4 = ec[3][0], Y4 = ec[3][1];
    double J11 = (X1 * (s - 1) / 4 - X2 * (s - 1) / 4 + X3 * (s + 1) / 4 - X4 * (s + 1) / 4);
    double J21 = (Y1 * (s - 1) / 4 - Y2 * (s - 1) / 4 + Y3 * (s + 1) / 4 - Y4 * (s + 1) / 4);
    double J12 = (X1 * (r - 1) / 4 - X2 * (r + 1) / 4 + X3 * (r + 1) / 4 - X4 * (r - 1) / 4);
    double J22 = (Y1 * (r - 1) / 4 - Y2 * (r + 1) / 4 + Y3 * (r + 1) / 4 - Y4 * (r - 1) / 4);
    double A = sqrt(J11 * J11 + J21 * J21);
    double B = sqrt(J12 * J12 + J22 * J22);
    double ca = J11 / A, sa = J21 / A, cb = J12 / B, sb = J22 / B;
    double detJ = J11 * J22 - J12 * J21;
    double Ax = X1 - X2 - X3 + X4, Ay = Y1 - Y2 - Y3 + Y4;
    double Bx = X1 - X2 + X3 - X4, By = Y1 - Y2 + Y3 - Y4;
    double Cx = X1 + X2 - X3 - X4, Cy = Y1 + Y2 - Y3 - Y4;
    double P = sqrt((Ax + Bx * s) * (Ax + Bx * s) + (Ay + By * s) * (Ay + By * s));
    double Q = sqrt((Bx * r + Cx) * (Bx * r + Cx) + (By * r + Cy) * (By * r + Cy));
    double *g1 = Bs, *g2 = Bs + 24;
    /* gxz */
    g1[2] += (sa * (r + 1) * P - sb * (s + 1) * Q) / (16 * detJ);
    g1[3] += (-sa * (Y1 - Y4) * (r + 1) * P + sb * (Y1 - Y2) * (s + 1) * Q) / (32 * detJ);
    g1[4] += (sa * (X1 - X4) * (r + 1) * P - sb * (X1 - X2) * (s + 1) * Q) / (32 * detJ);
    g1[8] += (-sa * (r - 1) * P + sb * (s + 1) * Q) / (16 * detJ);
    g1[9] += (sa * (Y2 - Y3) * (r - 1) * P + sb * (Y1 - Y2) * (s + 1) * Q) / (32 * detJ);
    g1[10] += (-sa * (X2 - X3) * (r - 1) * P - sb * (X1 - X2) * (s + 1) * Q) / (32 * detJ);
    g1[14] += (sa * (r - 1) * P - sb * (s - 1) * Q) / (16 * detJ);
    g1[15] += (sa * (Y2 - Y3) * (r - 1) * P + sb * (Y3 - Y4) * (s - 1) * Q) / (32 * detJ);
    g1[16] += (-sa * (X2 - X3) * (r - 1) * P - sb * (X3 - X4) * (s - 1) * Q) / (32 * detJ);
    g1[20] += (-sa * (r + 1) * P + sb * (s - 1) * Q) / (16 * detJ);
    g1[21] += (-sa * (Y1 - Y4) * (r + 1) * P + sb * (Y3 - Y4) * (s - 1) * Q) / (32 * detJ);
    g1[22] += (sa * (X1 - X4) * (r + 1) * P - sb * (X3 - X4) * (s - 1) * Q) / (32 * detJ);
    /* gyz */
    g2[2] += (-ca * (r + 1) * P + cb * (s + 1) * Q) / (16 * detJ);
    g2[3] += (ca * (Y1 - Y4) * (r + 1) * P - cb * (Y1 - Y2) * (s + 1) * Q) / (32 * detJ);
    g2[4] += (-ca * (X1 - X4) * (r + 1) * P + cb * (X1 - X2) * (s + 1) * Q) / (32 * detJ);
    g2[8] += (ca * (r - 1) * P - cb * (s + 1) * Q) / (16 * detJ);
    g2[9] += (-ca * (Y2 - Y3) * (r - 1) * P - cb * (Y1 - Y2) * (s + 1) * Q) / (32 * detJ);
    g2[10] += (ca * (X2 - X3) * (r - 1) * P + cb * (X1 - X2) * (s + 1) * Q) / (32 * detJ);
    g2[14] += (-ca * (r - 1) * P + cb * (s - 1) * Q) / (16 * detJ);
    g2[15] += (-ca * (Y2 - Y3) * (r - 1) * P - cb * (Y3 - Y4) * (s - 1) * Q) / (32 * detJ);
    g2[16] += (ca * (X2 - X3) * (r - 1) * P + cb * (X3 - X4) * (s - 1) * Q) / (32 * detJ);
    g2[20] += (ca * (r + 1) * P - cb * (s - 1) * Q) / (16 * detJ);
    g2[21] += (ca * (Y1 - Y4) * (r + 1) * P - cb * (Y3 - Y4) * (s - 1) * Q) / (32 * detJ);
    g2[22] += (-ca * (X1 - X4) * (r + 1) * P + cb * (X3 - X4) * (s - 1) * Q) / (32 * detJ);
}

/* B = Bloc * T for nstr x 24 matrices (mul!(Bm, tempBm, T)). */
static void bt(int nstr, const double *Bl, const double *T, double *B) {
    for (int i = 0; i < nstr; ++i)
        for (int j = 0; j < 24; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 24; ++k) acc += Bl[i * 24 + k] * T[k * 24 + j];
            B[i * 24 + j] = acc;
        }
}

/*
 * Q4RS element stiffness (FEMMShellQ4RSModule.jl:914-944) and Q4RSComp
 * (FEMMShellQ4RSCompModule.jl:903-955).  Arguments as for or_t3ff_elmat
 * with 4 nodes; Ke is 24 x 24.  Returns 0, or -1 for a singular metric
 * (FEMMShellQ4RSModule.jl:308-310).
 */
int or_q4rs_elmat(const double *X, const double *nrm, const uint8_t *valid, int composite,
                  const double *mat, const double *csys, double t, double stab_c, double drill_scale,
                  double *Ke) {
    const int kd = 24;
    double Dps[9], Dt[4], sA[9], sB[9], sD[9], sH[4];
    if (!composite) {
        memcpy(Dps, mat, 9 * sizeof(double));
        memcpy(Dt, mat + 9, 4 * sizeof(double));
        double scf = 5.0 / 6.0;
        for (int i = 0; i < 4; ++i) Dt[i] *= scf;
    }
    /* _quad_diameter :861-870 */
    double maxd = 0.0;
    for (int i = 1; i < 4; ++i) {
        double dx = X[i * 3 + 0] - X[0], dy = X[i * 3 + 1] - X[1], dz = X[i * 3 + 2] - X[2];
        double dist = (dx * dx + dy * dy + dz * dz);
        maxd = (dist > maxd) ? dist : maxd;
    }
    double h = sqrt(maxd);
    double stab = (t * t) / (t * t + stab_c * (h * h));
    memset(Ke, 0, sizeof(double) * kd * kd);
    for (int j = 0; j < 4; ++j) {
        double xi, eta, N[4], gN[4][2], J[3][2], EG[3][3];
        q4_gauss_point(j, &xi, &eta);
        q4_bfun(xi, eta, N, gN);
        q4_jac(X, gN, J);
        double Jac = surf_jac(J);
        double w = 1.0;
        e_g(EG, J);
        /* _EcoordsE :588-608 */
        double cen[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 3; ++k) cen[k] += X[a * 3 + k];
        for (int k = 0; k < 3; ++k) cen[k] /= 4;
        double ec[4][2];
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 2; ++k) {
                double d = 0.0;
                for (int m = 0; m < 3; ++m) d += (X[a * 3 + m] - cen[m]) * EG[m][k];
                ec[a][k] = d;
            }
        /* _LocalDerivatives :305-323 */
        double G[2][2];
        for (int p = 0; p < 2; ++p)
            for (int q = 0; q < 2; ++q) G[p][q] = J[0][p] * J[0][q] + J[1][p] * J[1][q] + J[2][p] * J[2][q];
        double detG = G[0][0] * G[1][1] - G[0][1] * G[1][0];
        if (detG == 0.0) return -1;
        double iG[2][2] = {{G[1][1] / detG, -G[0][1] / detG}, {-G[1][0] / detG, G[0][0] / detG}};
        double gradN[4][2];
        for (int a = 0; a < 4; ++a) {
            double t2[2] = {iG[0][0] * gN[a][0] + iG[0][1] * gN[a][1], iG[1][0] * gN[a][0] + iG[1][1] * gN[a][1]};
            double g[3];
            for (int i = 0; i < 3; ++i) g[i] = J[i][0] * t2[0] + J[i][1] * t2[1];
            gradN[a][0] = EG[0][0] * g[0] + EG[1][0] * g[1] + EG[2][0] * g[2];
            gradN[a][1] = EG[0][1] * g[0] + EG[1][1] * g[1] + EG[2][1] * g[2];
        }
        double AEs[4][3][3];
        int nvalid[4];
        double Tga[24 * 24], Tae[24 * 24], T[24 * 24];
        nodal_triads_e(4, AEs, nvalid, EG, nrm, valid);
        transfmat_g_to_a(4, Tga, AEs, EG);
        transfmat_a_to_e(4, Tae, AEs, &gradN[0][0]);
        for (int a = 0; a < 24; ++a)
            for (int b = 0; b < 24; ++b) {
                double acc = 0.0;
                for (int k = 0; k < 24; ++k) acc += Tae[a * 24 + k] * Tga[k * 24 + b];
                T[a * 24 + b] = acc;
            }
        if (composite) {
            double m, n;
            layup2element_angle(EG, (const double(*)[3])csys, &m, &n);
            rotate_laminate(m, n, mat, mat + 9, mat + 18, mat + 27, sA, sB, sD, sH);
        }
        double Bl[3 * 24], Bm[3 * 24], Bb[3 * 24], Bsl[2 * 24], Bs[2 * 24];
        memset(Bl, 0, sizeof(Bl));
        for (int i = 0; i < 4; ++i) {
            int off = i * NDOF;
            Bl[0 * 24 + off + 0] = gradN[i][0];
            Bl[1 * 24 + off + 1] = gradN[i][1];
            Bl[2 * 24 + off + 0] = gradN[i][1];
            Bl[2 * 24 + off + 1] = gradN[i][0];
        }
        bt(3, Bl, T, Bm);
        memset(Bl, 0, sizeof(Bl));
        for (int i = 0; i < 4; ++i) {
            int off = i * NDOF;
            Bl[0 * 24 + off + 4] = gradN[i][0];
            Bl[1 * 24 + off + 3] = -gradN[i][1];
            Bl[2 * 24 + off + 3] = -gradN[i][0];
            Bl[2 * 24 + off + 4] = gradN[i][1];
        }
        bt(3, Bl, T, Bb);
        q4_bsmat_loc(Bsl, ec, xi, eta);
        bt(2, Bsl, T, Bs);
        if (!composite) {
            add_btdb_ut_only(kd, Ke, 3, Bm, t * Jac * w, Dps);
            add_btdb_ut_only(kd, Ke, 3, Bb, (t * t * t / 12.0) * Jac * w, Dps);
            add_btdb_ut_only(kd, Ke, 2, Bs, t * stab * Jac * w, Dt);
        } else {
            add_btdb_ut_only(kd, Ke, 3, Bm, Jac * w, sA);
            add_btdb_ut_only(kd, Ke, 3, Bb, Jac * w, sD);
            add_b1tdb2(kd, Ke, 3, Bm, Bb, Jac * w, sB);
            add_b1tdb2(kd, Ke, 3, Bb, Bm, Jac * w, sB);
            add_btdb_ut_only(kd, Ke, 2, Bs, stab * Jac * w, sH);
        }
    }
    complete_lt(kd, Ke);
    /* _add_drilling_stiffness! :807-859 (global basis, stored normals) */
    if (drill_scale != 0.0) {
        double sum = 0.0;
        int cnt = 0;
        for (int k = 0; k < 4; ++k) {
            if (!valid[k]) continue;
            const double *nv = nrm + 3 * k;
            double nn = norm3(nv);
            if (nn == 0.0) continue;
            double u[3] = {nv[0] / nn, nv[1] / nn, nv[2] / nn};
            double P[9], Krr[9], tmp[9], Kt[9];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    P[a * 3 + b] = ((a == b) ? 1.0 : 0.0) - u[a] * u[b];
                    Krr[a * 3 + b] = Ke[(k * NDOF + 3 + a) * kd + k * NDOF + 3 + b];
                }
            matmul(3, P, Krr, tmp);
            matmul(3, tmp, P, Kt);
            double tr = ((Kt[0] + Kt[4]) + Kt[8]) / 2.0;
            sum += (tr > 0.0) ? tr : 0.0;
            cnt++;
        }
        double kavg = (cnt == 0) ? 0.0 : (sum / cnt) * drill_scale;
        if (kavg != 0.0) {
            for (int k = 0; k < 4; ++k) {
                if (!valid[k]) continue;
                const double *nv = nrm + 3 * k;
                if (norm3(nv) == 0.0) continue;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        Ke[(k * NDOF + 3 + a) * kd + k * NDOF + 3 + b] += kavg * (nv[a] * nv[b]);
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Nodal normals (associategeometry!)                                         */
/* ------------------------------------------------------------------------ */

/* T3 isoparametric normals, FEMMShellT3FFModule.jl:570-616.  xyz N x 3,
 * conn E x 3 (0-based).  normals must be zero on entry (constructor :200). */
void or_t3_normals(int64_t N, const double *xyz, int64_t E, const int64_t *conn, double threshold_deg,
                   double *normals, uint8_t *valid) {
    for (int64_t i = 0; i < N; ++i) valid[i] = 1;
    for (int64_t el = 0; el < E; ++el) {
        double X[9], J0[3][2], EG[3][3];
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) X[a * 3 + k] = xyz[conn[el * 3 + a] * 3 + k];
        for (int j = 0; j < 3; ++j) {
            J0[j][0] = X[3 + j] - X[j];
            J0[j][1] = X[6 + j] - X[j];
        }
        e_g(EG, J0);
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) normals[conn[el * 3 + a] * 3 + k] += EG[k][2];
    }
    for (int64_t i = 0; i < N; ++i) {
        double nn = norm3(normals + 3 * i);
        if (nn > 0.0)
            for (int k = 0; k < 3; ++k) normals[i * 3 + k] /= nn;
    }
    double s = sin(threshold_deg / 180 * M_PI);
    double ntol = 1 - sqrt(1 - s * s);
    for (int64_t el = 0; el < E; ++el) {
        double X[9], J0[3][2], EG[3][3];
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) X[a * 3 + k] = xyz[conn[el * 3 + a] * 3 + k];
        for (int j = 0; j < 3; ++j) {
            J0[j][0] = X[3 + j] - X[j];
            J0[j][1] = X[6 + j] - X[j];
        }
        e_g(EG, J0);
        for (int a = 0; a < 3; ++a) {
            const double *nv = normals + 3 * conn[el * 3 + a];
            double nd = nv[0] * EG[0][2] + nv[1] * EG[1][2] + nv[2] * EG[2][2];
            if (nd < 1 - ntol) valid[conn[el * 3 + a]] = 0;
        }
    }
}

/* Q4 isoparametric normals, FEMMShellQ4RSModule.jl:472-525: area-weighted
 * element normals at the nodes (NodalTensorProductRule(2)). */
void or_q4_normals(int64_t N, const double *xyz, int64_t E, const int64_t *conn, double threshold_deg,
                   double *normals, uint8_t *valid) {
    static const double npc[4][2] = {{-1.0, -1.0}, {1.0, -1.0}, {1.0, 1.0}, {-1.0, 1.0}};
    for (int64_t i = 0; i < N; ++i) {
        valid[i] = 1;
        normals[3 * i] = normals[3 * i + 1] = normals[3 * i + 2] = 0.0;
    }
    for (int pass = 0; pass < 2; ++pass) {
        double s = sin(threshold_deg / 180 * M_PI);
        double ntol = 1 - sqrt(1 - s * s);
        for (int64_t el = 0; el < E; ++el) {
            double X[12];
            for (int a = 0; a < 4; ++a)
                for (int k = 0; k < 3; ++k) X[a * 3 + k] = xyz[conn[el * 4 + a] * 3 + k];
            for (int j = 0; j < 4; ++j) {
                double Nn[4], gN[4][2], J[3][2], EG[3][3];
                q4_bfun(npc[j][0], npc[j][1], Nn, gN);
                q4_jac(X, gN, J);
                e_g(EG, J);
                int64_t n = conn[el * 4 + j];
                if (pass == 0) {
                    double Jac = surf_jac(J);
                    for (int k = 0; k < 3; ++k) normals[n * 3 + k] += Jac * EG[k][2];
                } else {
                    const double *nv = normals + 3 * n;
                    double nd = nv[0] * EG[0][2] + nv[1] * EG[1][2] + nv[2] * EG[2][2];
                    if (nd < 1 - ntol) valid[n] = 0;
                }
            }
        }
        if (pass == 0)
            for (int64_t i = 0; i < N; ++i) {
                double nn = norm3(normals + 3 * i);
                if (nn > 0.0)
                    for (int k = 0; k < 3; ++k) normals[i * 3 + k] /= nn;
            }
    }
}

/* ------------------------------------------------------------------------ */
/* Lumped mass                                                                */
/* ------------------------------------------------------------------------ */

/* Per-element translational and rotational lumped masses.
 * T3FF :781-805 (tmass = rho t Ae / 3, rmass = rho t^3/12 Ae / 3);
 * T3FFComp :738-763 (mass_density Ae / 3, moment_density Ae / 3);
 * Q4RS :989-1016 and Q4RSComp :1004-1016 (2x2 Gauss, divided by 4).
 * m0 = rho*t (mass per area), m1 = rho*t^3/12 (rotary inertia per area); for
 * composites m0, m1 are the laminate_inertia! values. */
void or_element_masses(int nen, int64_t E, const double *xyz, const int64_t *conn, double m0, double m1,
                       int homogeneous, double rho, double t, double *tm, double *rm) {
    for (int64_t el = 0; el < E; ++el) {
        double X[12];
        for (int a = 0; a < nen; ++a)
            for (int k = 0; k < 3; ++k) X[a * 3 + k] = xyz[conn[el * nen + a] * 3 + k];
        if (nen == 3) {
            double J0[3][2], EG[3][3], ec[3][2], gradN[3][2], Ae, cen[3];
            t3_geometry(X, J0, EG, ec, gradN, &Ae, cen);
            if (homogeneous) {
                tm[el] = rho * (t * Ae) / 3;
                rm[el] = rho * (t * t * t / 12 * Ae) / 3;
            } else {
                tm[el] = m0 * (Ae / 3);
                rm[el] = m1 * (Ae / 3);
            }
        } else {
            double a = 0.0, b = 0.0;
            for (int j = 0; j < 4; ++j) {
                double xi, eta, N[4], gN[4][2], J[3][2];
                q4_gauss_point(j, &xi, &eta);
                q4_bfun(xi, eta, N, gN);
                q4_jac(X, gN, J);
                double Jac = surf_jac(J);
                if (homogeneous) {
                    a += rho * t * Jac * 1.0;
                    b += rho * (t * t * t) / 12 * Jac * 1.0;
                } else {
                    a += m0 * Jac * 1.0;
                    b += m1 * Jac * 1.0;
                }
            }
            tm[el] = a / 4;
            rm[el] = b / 4;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Assembly: COO append in element order, duplicates summed in input order   */
/* (FinEtools SysmatAssemblerSparse -> sparse(I,J,V); AssemblyModule.jl:47-52) */
/* ------------------------------------------------------------------------ */

/*
 * Assemble E element matrices (E x k x k, row-major, k = 6 nen) into a CSR
 * matrix of size nrows x nrows.  dofnums: N x 6 row-major, 0-based global
 * equation numbers.  Entries whose row or column is >= nrows are dropped
 * (FFBlock when nrows = nfreedofs; all dofs when nrows = nalldofs).
 * The COO stream is the one assemble! appends (element order; within an
 * element column-major: for j in cols, for i in rows).  As in Julia's
 * sparse(I, J, V): a stable counting sort by row, then duplicates combined
 * in input order, then each row's columns sorted.
 * Call with nzval == NULL to get nnz (rowptr is filled).  rowptr has
 * nrows + 1 entries; colval/nzval at least nnz entries (or_assemble_count
 * gives an upper bound).  Returns nnz, or -1 on OOM.
 * The counting and bucketing passes run on threads that each own a range of
 * rows and scan the whole COO stream in order, so every row receives its
 * entries in input order exactly as a serial pass would.
 */
static void coo_rows_pass(int nen, int64_t E, const int64_t *conn, const int64_t *dofnums, int64_t nrows,
                          const double *Ke, int64_t *cnt, int64_t *fill, int64_t *bc, double *bv) {
    int k = NDOF * nen;
#pragma omp parallel
    {
        int nth = 1, tid = 0;
#ifdef _OPENMP
        nth = omp_get_num_threads();
        tid = omp_get_thread_num();
#endif
        const int64_t r0 = nrows * tid / nth, r1 = nrows * (tid + 1) / nth;
        for (int64_t el = 0; el < E; ++el) {
            int64_t dn[24];
            for (int a = 0; a < nen; ++a)
                for (int d = 0; d < NDOF; ++d) dn[a * NDOF + d] = dofnums[conn[el * nen + a] * NDOF + d];
            for (int j = 0; j < k; ++j)
                if (dn[j] < nrows)
                    for (int i = 0; i < k; ++i)
                        if (dn[i] >= r0 && dn[i] < r1) {
                            if (!fill) {
                                cnt[dn[i] + 1]++;
                            } else {
                                int64_t q = fill[dn[i]]++;
                                bc[q] = dn[j];
                                bv[q] = Ke ? Ke[el * k * k + i * k + j] : 0.0;
                            }
                        }
        }
    }
}

/* Number of COO entries of the block (an upper bound of its nnz). */
int64_t or_assemble_count(int nen, int64_t E, const int64_t *conn, const int64_t *dofnums, int64_t nrows) {
    int64_t *cnt = (int64_t *)calloc(nrows + 1, sizeof(int64_t));
    if (!cnt) return -1;
    coo_rows_pass(nen, E, conn, dofnums, nrows, NULL, cnt, NULL, NULL, NULL);
    int64_t m = 0;
    for (int64_t r = 0; r < nrows; ++r) m += cnt[r + 1];
    free(cnt);
    return m;
}

int64_t or_assemble_csr(int nen, int64_t E, const int64_t *conn, const int64_t *dofnums, int64_t nrows,
                        const double *Ke, int64_t *rowptr, int64_t *colval, double *nzval) {
    int64_t *cnt = (int64_t *)calloc(nrows + 1, sizeof(int64_t));
    if (!cnt) return -1;
    /* pass 1: entries per row */
    coo_rows_pass(nen, E, conn, dofnums, nrows, NULL, cnt, NULL, NULL, NULL);
    for (int64_t r = 0; r < nrows; ++r) cnt[r + 1] += cnt[r];
    int64_t m = cnt[nrows];
    int64_t *bc = (int64_t *)malloc(sizeof(int64_t) * (m > 0 ? m : 1));
    double *bv = (double *)malloc(sizeof(double) * (m > 0 ? m : 1));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (nrows + 1));
    if (!bc || !bv || !fill) {
        free(cnt); free(bc); free(bv); free(fill);
        return -1;
    }
    memcpy(fill, cnt, sizeof(int64_t) * (nrows + 1));
    /* pass 2: stable bucket by row (input order kept within a row) */
    coo_rows_pass(nen, E, conn, dofnums, nrows, Ke, cnt, fill, bc, bv);
    /* pass 3: per row, combine duplicates in input order, then sort columns */
    int64_t *rlen = (int64_t *)calloc(nrows + 1, sizeof(int64_t));
    int nth = 1;
#ifdef _OPENMP
    nth = omp_get_max_threads();
#endif
    int64_t *marks = (int64_t *)malloc(sizeof(int64_t) * (size_t)nth * (nrows > 0 ? nrows : 1));
    if (!rlen || !marks) {
        free(cnt); free(bc); free(bv); free(fill); free(rlen); free(marks);
        return -1;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)nth * nrows; ++i) marks[i] = -1;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t r = 0; r < nrows; ++r) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        int64_t *mk = marks + (int64_t)tid * nrows;
        int64_t b = cnt[r], e = cnt[r + 1], w = b;
        for (int64_t q = b; q < e; ++q) {
            int64_t c = bc[q];
            if (mk[c] >= b && mk[c] < w && bc[mk[c]] == c) {
                bv[mk[c]] = bv[mk[c]] + bv[q];
            } else {
                mk[c] = w;
                bc[w] = c;
                bv[w] = bv[q];
                ++w;
            }
        }
        /* insertion sort of the compacted row by column */
        for (int64_t q = b + 1; q < w; ++q) {
            int64_t c = bc[q];
            double v = bv[q];
            int64_t p = q - 1;
            while (p >= b && bc[p] > c) {
                bc[p + 1] = bc[p];
                bv[p + 1] = bv[p];
                --p;
            }
            bc[p + 1] = c;
            bv[p + 1] = v;
        }
        for (int64_t q = b; q < w; ++q) mk[bc[q]] = -1;
        rlen[r + 1] = w - b;
    }
    rowptr[0] = 0;
    for (int64_t r = 0; r < nrows; ++r) rowptr[r + 1] = rowptr[r] + rlen[r + 1];
    int64_t nnz = rowptr[nrows];
    if (nzval) {
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < nrows; ++r) {
            int64_t len = rowptr[r + 1] - rowptr[r];
            memcpy(colval + rowptr[r], bc + cnt[r], sizeof(int64_t) * len);
            memcpy(nzval + rowptr[r], bv + cnt[r], sizeof(double) * len);
        }
    }
    free(cnt); free(bc); free(bv); free(fill); free(rlen); free(marks);
    return nnz;
}

/* ------------------------------------------------------------------------ */
/* Explicit central-difference loop (plate_expl_examples.jl:61-94)            */
/* ------------------------------------------------------------------------ */

static void csr_mv(int64_t n, const int64_t *rp, const int64_t *ci, const double *v, const double *x,
                   double *y) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n; ++r) {
        double acc = 0.0;
        for (int64_t p = rp[r]; p < rp[r + 1]; ++p) acc += v[p] * x[ci[p]];
        y[r] = acc;
    }
}

/*
 * Runs nsteps of the central-difference loop with mass-proportional damping
 * C = cdamp * diag(M).  The external force is F(t_n) = g[n] * Fmag, with g
 * sampled by the caller at t_n = n dt (n = 0 .. nsteps).  dt is used as
 * given (the caller performs the tend/dt adjustment of :72-75).  invMC is
 * formed with dt_mass: the reference computes invMC (:70-71) from the dt it
 * was called with, before the adjustment of :72-75 can change dt, so a
 * damped loop whose dt was adjusted uses two step sizes.  U, V are the
 * initial conditions on entry and the final state on exit.  probe_out
 * (nsteps+1) x nprobe receives U[probes] after each step (peek).
 */
void or_cd_loop(int64_t n, const int64_t *rp, const int64_t *ci, const double *kv, const double *Mdiag,
                double cdamp, double dt, double dt_mass, int64_t nsteps, const double *Fmag, const double *g,
                double *U, double *V, int nprobe, const int64_t *probes, double *probe_out) {
    double *A = (double *)malloc(sizeof(double) * n);
    double *F = (double *)malloc(sizeof(double) * n);
    double *Ev = (double *)malloc(sizeof(double) * n);
    double *C = (double *)malloc(sizeof(double) * n);
    double *invMC = (double *)malloc(sizeof(double) * n);
    for (int64_t i = 0; i < n; ++i) {
        C[i] = cdamp * Mdiag[i];
        invMC[i] = 1.0 / (Mdiag[i] + (dt_mass / 2) * C[i]);
    }
    double dt2_2 = ((dt * dt) / 2), dt_2 = (dt / 2);
    for (int64_t i = 0; i < n; ++i) {
        F[i] = g[0] * Fmag[i];
        A[i] = invMC[i] * F[i];
    }
    for (int p = 0; p < nprobe; ++p) probe_out[p] = U[probes[p]];
    for (int64_t step = 1; step <= nsteps; ++step) {
        for (int64_t i = 0; i < n; ++i) U[i] += dt * V[i] + dt2_2 * A[i];
        for (int64_t i = 0; i < n; ++i) F[i] = g[step] * Fmag[i];
        csr_mv(n, rp, ci, kv, U, Ev);
        for (int64_t i = 0; i < n; ++i) F[i] -= Ev[i] + C[i] * (V[i] + dt_2 * A[i]);
        for (int64_t i = 0; i < n; ++i) V[i] += dt_2 * A[i];
        for (int64_t i = 0; i < n; ++i) A[i] = invMC[i] * F[i];
        for (int64_t i = 0; i < n; ++i) V[i] += dt_2 * A[i];
        for (int p = 0; p < nprobe; ++p) probe_out[step * nprobe + p] = U[probes[p]];
    }
    free(A); free(F); free(Ev); free(C); free(invMC);
}

/*
 * Power iteration for the largest generalized eigenvalue of (K, M), the form
 * of pwr (clamp_cyl_transient_examples.jl:45-60), started from the given v
 * (the reference starts from rand(); a seeded start is used here).  Returns
 * omega_max = sqrt(w'v / v'Mv).
 */
double or_pwr(int64_t n, const int64_t *rp, const int64_t *ci, const double *kv, const double *Mdiag,
              double *v, int maxit) {
    double *w = (double *)malloc(sizeof(double) * n);
    for (int it = 0; it < maxit; ++it) {
        csr_mv(n, rp, ci, kv, v, w);
        double wn = 0.0;
        for (int64_t i = 0; i < n; ++i) wn += w[i] * w[i];
        wn = sqrt(wn);
        for (int64_t i = 0; i < n; ++i) w[i] *= (1.0 / wn);
        double vn = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            v[i] = (1.0 / Mdiag[i]) * w[i];
            vn += v[i] * v[i];
        }
        vn = sqrt(vn);
        for (int64_t i = 0; i < n; ++i) v[i] *= (1.0 / vn);
    }
    csr_mv(n, rp, ci, kv, v, w);
    double num = 0.0, den = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        num += w[i] * v[i];
        den += v[i] * (Mdiag[i] * v[i]);
    }
    free(w);
    return sqrt(num / den);
}

/* ------------------------------------------------------------------------ */
/* Composite layup (CompositeLayupModule.jl)                                  */
/* ------------------------------------------------------------------------ */

/* Ply constructor (:71-97): plane-stress Dps (3x3) and transverse Dts (2x2)
 * from the 6x6 3-D stiffness D (ordering 11,22,33,12,13,23). */
void or_ply_matrices(const double *D6, double *Dps, double *Dts) {
    double d11 = D6[0 * 6 + 0] - (D6[0 * 6 + 2] * D6[2 * 6 + 0]) / D6[2 * 6 + 2];
    double d12 = D6[0 * 6 + 1] - (D6[0 * 6 + 2] * D6[2 * 6 + 1]) / D6[2 * 6 + 2];
    double d21 = D6[1 * 6 + 0] - (D6[1 * 6 + 2] * D6[2 * 6 + 0]) / D6[2 * 6 + 2];
    double d22 = D6[1 * 6 + 1] - (D6[1 * 6 + 2] * D6[2 * 6 + 1]) / D6[2 * 6 + 2];
    Dps[0] = d11; Dps[1] = d12; Dps[3] = d21; Dps[4] = d22;
    int ix[3] = {0, 1, 3};
    for (int i = 0; i < 3; ++i) {
        Dps[2 * 3 + i] = D6[3 * 6 + ix[i]];
        Dps[i * 3 + 2] = D6[3 * 6 + ix[i]];
    }
    Dts[0] = D6[4 * 6 + 4]; Dts[1] = 0.0; Dts[2] = 0.0; Dts[3] = D6[5 * 6 + 5];
}

/*
 * laminate_stiffnesses! (:246-271), laminate_transverse_stiffness! (:280-308,
 * Vinson-Sierakowski when tsc <= 0, else constant tsc) and laminate_inertia!
 * (:317-330).  Plies bottom to top; angle in degrees.
 */
void or_laminate(int np, const double *Dps, const double *Dts, const double *thick, const double *angle_deg,
                 const double *rho, double offset, double tsc, double *A, double *B, double *D, double *H,
                 double *inertia) {
    double tl = 0.0;
    for (int p = 0; p < np; ++p) tl += thick[p];
    for (int i = 0; i < 9; ++i) A[i] = B[i] = D[i] = 0.0;
    for (int i = 0; i < 4; ++i) H[i] = 0.0;
    double zs = -tl / 2 - offset;
    for (int p = 0; p < np; ++p) {
        double ze = zs + thick[p];
        double ang = angle_deg[p] / 180 * M_PI;
        double m = cos(-ang), n = sin(-ang);
        double T[9] = {m * m, n * n, 2 * (m * n), n * n, m * m, -2 * (m * n), -(m * n), (m * n), m * m - n * n};
        /* TransformerQEQt: Dps <- T * Dps * T' */
        double buf[9], Dr[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double acc = 0.0;
                for (int k = 0; k < 3; ++k) acc += Dps[p * 9 + i * 3 + k] * T[j * 3 + k];
                buf[i * 3 + j] = acc;
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double acc = 0.0;
                for (int k = 0; k < 3; ++k) acc += T[i * 3 + k] * buf[k * 3 + j];
                Dr[i * 3 + j] = acc;
            }
        for (int i = 0; i < 9; ++i) {
            A[i] += (ze - zs) * Dr[i];
            B[i] += (ze * ze - zs * zs) / 2 * Dr[i];
            D[i] += (ze * ze * ze - zs * zs * zs) / 3 * Dr[i];
        }
        zs += thick[p];
    }
    zs = -tl / 2 - offset;
    for (int p = 0; p < np; ++p) {
        double ze = zs + thick[p];
        double ang = angle_deg[p] / 180 * M_PI;
        double m = cos(ang), n = sin(ang);
        double Tt[4] = {m, -n, n, m};
        double Dt[4];
        memcpy(Dt, Dts + 4 * p, 4 * sizeof(double));
        qteq(2, Dt, Tt);
        for (int i = 0; i < 4; ++i) {
            if (tsc <= 0.0)
                H[i] += 5.0 / 4 * (ze - zs - 4.0 / 3 * (ze * ze * ze - zs * zs * zs) / (tl * tl)) * Dt[i];
            else
                H[i] += tsc * (ze - zs) * Dt[i];
        }
        zs += thick[p];
    }
    zs = -tl / 2 - offset;
    double md = 0.0, mi = 0.0;
    for (int p = 0; p < np; ++p) {
        double ze = zs + thick[p];
        md += (ze - zs) * rho[p];
        mi += (ze * ze * ze - zs * zs * zs) * rho[p] / 3;
        zs += thick[p];
    }
    inertia[0] = md;
    inertia[1] = mi;
}

/* ------------------------------------------------------------------------ */
/* Batch driver: the element loop of stiffness() for all elements             */
/* ------------------------------------------------------------------------ */

/*
 * Element stiffness matrices for all E elements (the loop bodies of
 * FEMMShellT3FFModule.jl:670-734, FEMMShellT3FFCompModule.jl:606-686,
 * FEMMShellQ4RSModule.jl:914-945, FEMMShellQ4RSCompModule.jl:903-956).
 * nen = 3 (T3FF family) or 4 (Q4RS family).  xyz: N x 3; conn: E x nen
 * (0-based); normals N x 3; valid N.  For composites, mat holds ngroups
 * blocks of 31 doubles (A, B, D, H), csys ngroups blocks of 9, tgroup the
 * layup thickness per group and elem_group the group of each element (NULL
 * = group 0).  Ke: E x k x k.  Returns 0 or -1 (singular Q4 metric).
 */
int or_stiffness_batch(int nen, int64_t E, const double *xyz, const int64_t *conn, const double *normals,
                       const uint8_t *valid, int composite, const double *mat, const double *csys,
                       const double *tgroup, const int32_t *elem_group, double t, double stab_c,
                       int formulation, double drill_scale, double *Ke) {
    int k = NDOF * nen;
    int status = 0;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t el = 0; el < E; ++el) {
        double X[12], nrm[12];
        uint8_t vl[4];
        for (int a = 0; a < nen; ++a) {
            int64_t n = conn[el * nen + a];
            for (int c = 0; c < 3; ++c) {
                X[a * 3 + c] = xyz[n * 3 + c];
                nrm[a * 3 + c] = normals[n * 3 + c];
            }
            vl[a] = valid[n];
        }
        int g = (composite && elem_group) ? elem_group[el] : 0;
        const double *m = composite ? mat + 31 * g : mat;
        const double *cs = composite ? csys + 9 * g : csys;
        double tt = composite ? tgroup[g] : t;
        if (nen == 3) {
            or_t3ff_elmat(X, nrm, vl, composite, m, cs, tt, stab_c, formulation, drill_scale,
                          Ke + el * (int64_t)k * k);
        } else {
            if (or_q4rs_elmat(X, nrm, vl, composite, m, cs, tt, stab_c, drill_scale, Ke + el * (int64_t)k * k))
                status = -1;
        }
    }
    return status;
}

/* ------------------------------------------------------------------------ */
/* Stress resultants (inspectintegpoints)                                     */
/* ------------------------------------------------------------------------ */

static void matvec(int n, const double *A, const double *x, double *y) {
    for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int k = 0; k < n; ++k) acc += A[i * n + k] * x[k];
        y[i] = acc;
    }
}

static void rectvec(int m, int n, const double *A, const double *x, double *y) {
    for (int i = 0; i < m; ++i) {
        double acc = 0.0;
        for (int k = 0; k < n; ++k) acc += A[i * n + k] * x[k];
        y[i] = acc;
    }
}

/* o2_e' * [a1 a3; a3 a2] * o2_e -> (X11, X22, X12), evaluated as (o2' m) o2. */
static void rotate_tensor(const double o2[2][2], const double *a, double *out) {
    double m[2][2] = {{a[0], a[2]}, {a[2], a[1]}}, t[2][2], X[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) t[i][j] = o2[0][i] * m[0][j] + o2[1][i] * m[1][j];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) X[i][j] = t[i][0] * o2[0][j] + t[i][1] * o2[1][j];
    out[0] = X[0][0];
    out[1] = X[1][1];
    out[2] = X[0][1];
}

/* Material law and output frame of one integration point.  Homogeneous
 * (FEMMShellT3FFModule.jl:929-957, FEMMShellQ4RSModule.jl:1139-1163):
 * N = (t Dps) e, M = (t^3/12 Dps) k, Q = (t stab (5/6 Dt)) g.  Layered
 * (FEMMShellT3FFCompModule.jl:891-937, FEMMShellQ4RSCompModule.jl:1157-1201):
 * A, B, D, H rotated into the element frame by the layup angle, N = A e + B k,
 * M = B e + D k, Q = (stab H) g.  Then the tensors are rotated into the
 * output csys with o2_e = [m n; -n m] from Layup2ElementAngle(E_G, ocs); a
 * NULL ocs is the isoparametric csys, which is E_G itself. */
static void point_resultants(int composite, const double *mat, const double *lcs, const double EG[3][3], double t,
                             double stab, const double *ocs, const double *eps, const double *kap,
                             const double *gam, double *out) {
    double f[3], mo[3], q[2];
    if (!composite) {
        const double *Dps = mat;
        double Dt[4];
        for (int i = 0; i < 4; ++i) Dt[i] = mat[9 + i] * (5.0 / 6.0);
        double cb = (t * t * t) / 12;
        double cs = t * stab;
        for (int i = 0; i < 3; ++i) {
            double a = 0.0, b = 0.0;
            for (int j = 0; j < 3; ++j) {
                a += (t * Dps[i * 3 + j]) * eps[j];
                b += (cb * Dps[i * 3 + j]) * kap[j];
            }
            f[i] = a;
            mo[i] = b;
        }
        for (int i = 0; i < 2; ++i) q[i] = (cs * Dt[i * 2 + 0]) * gam[0] + (cs * Dt[i * 2 + 1]) * gam[1];
    } else {
        double m, n, sA[9], sB[9], sD[9], sH[4];
        layup2element_angle(EG, (const double(*)[3])lcs, &m, &n);
        rotate_laminate(m, n, mat, mat + 9, mat + 18, mat + 27, sA, sB, sD, sH);
        for (int i = 0; i < 3; ++i) {
            double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
            for (int j = 0; j < 3; ++j) {
                a += sA[i * 3 + j] * eps[j];
                b += sB[i * 3 + j] * kap[j];
                c += sB[i * 3 + j] * eps[j];
                d += sD[i * 3 + j] * kap[j];
            }
            f[i] = a + b;
            mo[i] = c + d;
        }
        for (int i = 0; i < 2; ++i) q[i] = (stab * sH[i * 2 + 0]) * gam[0] + (stab * sH[i * 2 + 1]) * gam[1];
    }
    double om, on;
    if (ocs)
        layup2element_angle(EG, (const double(*)[3])ocs, &om, &on);
    else
        layup2element_angle(EG, EG, &om, &on);
    double o2[2][2] = {{om, on}, {-on, om}};
    rotate_tensor(o2, mo, out + 3);
    rotate_tensor(o2, f, out + 0);
    out[6] = o2[0][0] * q[0] + o2[1][0] * q[1];
    out[7] = o2[0][1] * q[0] + o2[1][1] * q[1];
}

/*
 * T3FF / T3FFComp inspectintegpoints (FEMMShellT3FFModule.jl:850-962,
 * FEMMShellT3FFCompModule.jl:809-943) at the one point (the centroid),
 * all three quantities: out[8] = N11 N22 N12 M11 M22 M12 Q1 Q2.
 *   ue: 18 element displacements and rotations, global basis.
 *   edisp_n = Tga ue, edisp_e = Tae edisp_n; e = Bm edisp_e, k = Bb edisp_e,
 *   g = Bs edisp_e (_Bsmat!, the average of the three orderings), h = sqrt(2 Ae).
 */
void or_t3ff_resultants(const double *X, const double *nrm, const uint8_t *valid, int composite,
                        const double *mat, const double *lcs, double t, double stab_c, const double *ue,
                        const double *ocs, double *out) {
    double J0[3][2], EG[3][3], ec[3][2], gradN[3][2], Ae, centroid[3];
    t3_geometry(X, J0, EG, ec, gradN, &Ae, centroid);
    double AEs[3][3][3];
    int nvalid[3];
    double T[18 * 18], en[18], ee[18];
    nodal_triads_e(3, AEs, nvalid, EG, nrm, valid);
    transfmat_g_to_a(3, T, AEs, EG);
    matvec(18, T, ue, en);
    transfmat_a_to_e(3, T, AEs, &gradN[0][0]);
    matvec(18, T, en, ee);
    double Bm[3 * 18], Bb[3 * 18], Bs[2 * 18];
    memset(Bm, 0, sizeof(Bm));
    memset(Bb, 0, sizeof(Bb));
    memset(Bs, 0, sizeof(Bs));
    for (int i = 0; i < 3; ++i) {
        Bm[0 * 18 + 6 * i + 0] = gradN[i][0];
        Bm[1 * 18 + 6 * i + 1] = gradN[i][1];
        Bm[2 * 18 + 6 * i + 0] = gradN[i][1];
        Bm[2 * 18 + 6 * i + 1] = gradN[i][0];
        Bb[0 * 18 + 6 * i + 4] = gradN[i][0];
        Bb[1 * 18 + 6 * i + 3] = -gradN[i][1];
        Bb[2 * 18 + 6 * i + 3] = -gradN[i][0];
        Bb[2 * 18 + 6 * i + 4] = gradN[i][1];
    }
    add_bsmat_o(Bs, ec, Ae, 0, 1, 2);
    add_bsmat_o(Bs, ec, Ae, 1, 2, 0);
    add_bsmat_o(Bs, ec, Ae, 2, 0, 1);
    for (int i = 0; i < 36; ++i) Bs[i] *= (1.0 / 3);
    double eps[3], kap[3], gam[2];
    rectvec(3, 18, Bm, ee, eps);
    rectvec(3, 18, Bb, ee, kap);
    rectvec(2, 18, Bs, ee, gam);
    double h = sqrt(2 * Ae);
    double stab = (t * t) / (t * t + stab_c * (h * h));
    point_resultants(composite, mat, lcs, EG, t, stab, ocs, eps, kap, gam, out);
}

/*
 * Q4RS / Q4RSComp inspectintegpoints (FEMMShellQ4RSModule.jl:1061-1170,
 * FEMMShellQ4RSCompModule.jl:1077-1209) at the four Gauss points:
 * out[4][8] as above, ocs: NULL or 4 x 9 (one csys matrix per point).
 *   T = Tae Tga; e = (Bm_loc T) ue, k = (Bb_loc T) ue, g = (Bs_loc T) ue.
 *   twice != 0 restates the layered form, which multiplies the already
 *   transformed edisp_e = T ue by B_loc T again (:1163, :1168-1170, :1196).
 */
int or_q4rs_resultants(const double *X, const double *nrm, const uint8_t *valid, int composite,
                       const double *mat, const double *lcs, double t, double stab_c, const double *ue,
                       const double *ocs, int twice, double *out) {
    double maxd = 0.0;
    for (int i = 1; i < 4; ++i) {
        double dx = X[i * 3 + 0] - X[0], dy = X[i * 3 + 1] - X[1], dz = X[i * 3 + 2] - X[2];
        double dist = (dx * dx + dy * dy + dz * dz);
        maxd = (dist > maxd) ? dist : maxd;
    }
    double h = sqrt(maxd);
    double stab = (t * t) / (t * t + stab_c * (h * h));
    for (int j = 0; j < 4; ++j) {
        double xi, eta, N[4], gN[4][2], J[3][2], EG[3][3];
        q4_gauss_point(j, &xi, &eta);
        q4_bfun(xi, eta, N, gN);
        q4_jac(X, gN, J);
        e_g(EG, J);
        double cen[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 3; ++k) cen[k] += X[a * 3 + k];
        for (int k = 0; k < 3; ++k) cen[k] /= 4;
        double ec[4][2];
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 2; ++k) {
                double d = 0.0;
                for (int m = 0; m < 3; ++m) d += (X[a * 3 + m] - cen[m]) * EG[m][k];
                ec[a][k] = d;
            }
        double G[2][2];
        for (int p = 0; p < 2; ++p)
            for (int q = 0; q < 2; ++q) G[p][q] = J[0][p] * J[0][q] + J[1][p] * J[1][q] + J[2][p] * J[2][q];
        double detG = G[0][0] * G[1][1] - G[0][1] * G[1][0];
        if (detG == 0.0) return -1;
        double iG[2][2] = {{G[1][1] / detG, -G[0][1] / detG}, {-G[1][0] / detG, G[0][0] / detG}};
        double gradN[4][2];
        for (int a = 0; a < 4; ++a) {
            double t2[2] = {iG[0][0] * gN[a][0] + iG[0][1] * gN[a][1], iG[1][0] * gN[a][0] + iG[1][1] * gN[a][1]};
            double g[3];
            for (int i = 0; i < 3; ++i) g[i] = J[i][0] * t2[0] + J[i][1] * t2[1];
            gradN[a][0] = EG[0][0] * g[0] + EG[1][0] * g[1] + EG[2][0] * g[2];
            gradN[a][1] = EG[0][1] * g[0] + EG[1][1] * g[1] + EG[2][1] * g[2];
        }
        double AEs[4][3][3];
        int nvalid[4];
        double Tga[24 * 24], Tae[24 * 24], T[24 * 24];
        nodal_triads_e(4, AEs, nvalid, EG, nrm, valid);
        transfmat_g_to_a(4, Tga, AEs, EG);
        transfmat_a_to_e(4, Tae, AEs, &gradN[0][0]);
        for (int a = 0; a < 24; ++a)
            for (int b = 0; b < 24; ++b) {
                double acc = 0.0;
                for (int k = 0; k < 24; ++k) acc += Tae[a * 24 + k] * Tga[k * 24 + b];
                T[a * 24 + b] = acc;
            }
        double Bl[3 * 24], Bm[3 * 24], Bb[3 * 24], Bsl[2 * 24], Bs[2 * 24];
        memset(Bl, 0, sizeof(Bl));
        for (int i = 0; i < 4; ++i) {
            int off = i * NDOF;
            Bl[0 * 24 + off + 0] = gradN[i][0];
            Bl[1 * 24 + off + 1] = gradN[i][1];
            Bl[2 * 24 + off + 0] = gradN[i][1];
            Bl[2 * 24 + off + 1] = gradN[i][0];
        }
        bt(3, Bl, T, Bm);
        memset(Bl, 0, sizeof(Bl));
        for (int i = 0; i < 4; ++i) {
            int off = i * NDOF;
            Bl[0 * 24 + off + 4] = gradN[i][0];
            Bl[1 * 24 + off + 3] = -gradN[i][1];
            Bl[2 * 24 + off + 3] = -gradN[i][0];
            Bl[2 * 24 + off + 4] = gradN[i][1];
        }
        bt(3, Bl, T, Bb);
        q4_bsmat_loc(Bsl, ec, xi, eta);
        bt(2, Bsl, T, Bs);
        double ud[24];
        if (twice)
            matvec(24, T, ue, ud);
        else
            memcpy(ud, ue, sizeof(ud));
        double eps[3], kap[3], gam[2];
        rectvec(3, 24, Bm, ud, eps);
        rectvec(3, 24, Bb, ud, kap);
        rectvec(2, 24, Bs, ud, gam);
        point_resultants(composite, mat, lcs, EG, t, stab, ocs ? ocs + 9 * j : NULL, eps, kap, gam, out + 8 * j);
    }
    return 0;
}

/* Resultants of the listed elements.  u: N x 6 nodal values (global basis);
 * felist: nlist 0-based element numbers; ocs_kind 0: isoparametric csys,
 * 1: one 3 x 3 matrix (row-major) for every point, 2: nlist x npts x 9;
 * out: nlist x npts x 8.  Material arguments as for or_stiffness_batch. */
int or_resultants_batch(int nen, const double *xyz, const int64_t *conn, const double *normals,
                        const uint8_t *valid, int composite, const double *mat, const double *csys,
                        const double *tgroup, const int32_t *elem_group, double t, double stab_c, const double *u,
                        int64_t nlist, const int64_t *felist, int ocs_kind, const double *ocs, int twice,
                        double *out) {
    int npts = nen == 3 ? 1 : 4;
    int status = 0;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t l = 0; l < nlist; ++l) {
        int64_t el = felist[l];
        double X[12], nrm[12], ue[24];
        uint8_t vl[4];
        for (int a = 0; a < nen; ++a) {
            int64_t n = conn[el * nen + a];
            for (int c = 0; c < 3; ++c) {
                X[a * 3 + c] = xyz[n * 3 + c];
                nrm[a * 3 + c] = normals[n * 3 + c];
            }
            for (int c = 0; c < NDOF; ++c) ue[a * NDOF + c] = u[n * NDOF + c];
            vl[a] = valid[n];
        }
        int g = elem_group ? elem_group[el] : 0;
        const double *m = mat + (composite ? 31 : 13) * g;
        const double *cs = composite ? csys + 9 * g : NULL;
        double tt = tgroup ? tgroup[g] : t;
        const double *oc = ocs_kind == 0 ? NULL : ocs_kind == 1 ? ocs : ocs + l * npts * 9;
        double *o = out + l * npts * 8;
        if (nen == 3) {
            or_t3ff_resultants(X, nrm, vl, composite, m, cs, tt, stab_c, ue, oc, o);
        } else {
            double oc4[36];
            if (ocs_kind == 1)
                for (int j = 0; j < 4; ++j) memcpy(oc4 + 9 * j, ocs, 9 * sizeof(double));
            if (or_q4rs_resultants(X, nrm, vl, composite, m, cs, tt, stab_c, ue,
                                   ocs_kind == 0 ? NULL : ocs_kind == 1 ? oc4 : oc, twice, o))
                status = -1;
        }
    }
    return status;
}
