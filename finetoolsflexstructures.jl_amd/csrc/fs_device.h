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
    // homogeneous material per group (ngroups x 16): chol(Dps) 9 | chol(5/6 Dt) 4 |
    // thickness | stab_c | drilling scale
    const double *g_homog;
};

// Scatter plan of one block (ALL or FF) -- see scatter_kernel.
// Node I owns the CSR rows dof(I, r); all of them share one column list made
// of the runs of I's neighbour nodes J (sorted).  For each neighbour slot k in
// [nbr_ptr[I], nbr_ptr[I+1]): nbr_pair[k] = (upper pair index << 1) | side
// (side 1: I is the pair's second node), nbr_off[2k], nbr_off[2k+1] = offset
// of J's free / fixed column run within I's rows (-1 = none).
// node_code[J*6+c]: 0..5 offset of comp c in J's free run, 8..13 in the fixed
// run, -1 not in the block.  node_rowstart[I*6+r]: CSR position of row
// dof(I, r), -1 if that row is not in the block.
struct ScatterPlan {
    int64_t nnodes;
    const int32_t *nbr_ptr;       // N + 1
    const int32_t *nbr_node;      // nbr_ptr[N]: neighbour node J
    const int32_t *nbr_pair;      // nbr_ptr[N]
    const int32_t *nbr_off;       // 2 * nbr_ptr[N]
    const int8_t *node_code;      // N x 6
    const int32_t *node_rowstart; // N x 6
    const int32_t *contrib_ptr;   // npairs + 1
    const int32_t *contrib;       // (block index << 1) | transposed, element order per pair
};

}  // namespace fs
