// fs_device.h -- data layouts shared by the host planner and the gfx950 kernels.
//
// Device-resident model layout (all structure-of-arrays, coalesced by element
// or node index):
//   x[N], y[N], z[N]                node coordinates (FP64)
//   conn[nen][E]                    int32 node numbers, 0-based, one plane per
//                                   local node so that element e reads conn[a*E+e]
//   nx[N], ny[N], nz[N], valid[N]   nodal normals and validity flags
//   Kb[E][nblk][36]                 element node-pair blocks (upper pairs a <= b,
//                                   row-major 6x6, global basis)
//   pair tables                     see ScatterPlan
#pragma once
#include <cstdint>

namespace fs {

// Parameters of one element-stiffness launch.  Material factors are Cholesky
// factors (upper triangular, row-major) so that K_e = sum C^T C with
// C = L * B * T; see the element kernel for the derivation.
struct ElemParams {
    int32_t nen;          // 3 or 4
    int32_t composite;    // 0 or 1
    int32_t formulation;  // 0 AVERAGE_B, 1 AVERAGE_K (T3)
    int32_t nelem;
    int64_t nnodes;
    double thickness;     // homogeneous thickness
    double stab_c;
    double drill_scale;
    double Lps[9];        // chol(Dps), upper, row-major
    double Lt[4];         // chol(5/6 Dt), upper, row-major
    const double *g_Labd; // ngroups x 36: chol([A B; B D]), upper, row-major
    const double *g_Lh;   // ngroups x 4: chol(H), upper, row-major
    const double *g_csys; // ngroups x 9: layup csys, row-major [i][j]
    const double *g_t;    // ngroups: layup thickness
    const int32_t *elem_group;  // nelem or nullptr
};

// Scatter plan of one block (ALL or FF) -- see fs_scatter_kernel.
struct ScatterPlan {
    int64_t npairs;            // upper node pairs (I <= J)
    const int32_t *pair_I;     // npairs
    const int32_t *pair_J;     // npairs
    const int32_t *pair_off;   // npairs x 4: offset of J's free run / fixed run in
                               // I's rows, then of I's runs in J's rows (-1 = none)
    const int32_t *contrib_ptr;  // npairs + 1
    const int32_t *contrib;      // (block index << 1) | transposed
    const int64_t *node_rowstart; // N x 6: CSR position of row dof(I,r), -1 = not in block
    const int8_t *node_code;     // N x 6: 0..5 free-run offset, 8..13 fixed-run, -1 excluded
};

}  // namespace fs
