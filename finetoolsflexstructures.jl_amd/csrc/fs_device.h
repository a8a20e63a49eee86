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
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fs {

// Rows of the largest free-free block the allocation-free entry points take:
// their per-device workspace holds three vectors of this length, 512 dot
// partials and 8 scalars (24 MiB + 4 KiB of FP64).
constexpr int64_t NOALLOC_MAX_ROWS = int64_t(1) << 20;
constexpr int64_t NOALLOC_DOUBLES = 3 * NOALLOC_MAX_ROWS + 512 + 8;

// Parameters of one element-stiffness launch.  Material factors are Cholesky
// factors (upper triangular, row-major) so that K_e = sum C^T C with
// C = L * B * T; see the element kernel for the derivation.
struct ElemParams {
    int32_t nen;          // 3 or 4
    int32_t composite;    // 0 or 1
    int32_t formulation;  // 0 AVERAGE_B, 1 AVERAGE_K (T3)
    int32_t nelem;        // elements of the mesh (connectivity stride)
    int32_t ebegin, eend; // element range of this launch, [ebegin, eend)
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
    // Split output for the mixed-precision operator (both nullptr: full blocks).
    // With split_cs set, Kb receives the membrane + bending part only, the
    // transverse-shear rows of C (global basis, already scaled by the material
    // factor, weight and stabilisation) go to split_cs[e][nsr][6 nen], and the
    // drilling term of each element node goes to split_dr[e][a][9] (3 x 3
    // rotation block, global basis, row-major; zeros where the reference adds
    // nothing).  nsr = 2 (T3 AVERAGE_B), 6 (T3 AVERAGE_K), 8 (Q4: 2 per Gauss
    // point, Gauss point major).
    double *split_cs;
    double *split_dr;
    // Fused assembly (TilePlan): workgroup b computes the elements
    // tile_elem[b NE .. b NE + NE - 1] (element numbers; a tile's unused
    // slots repeat one of its elements) instead of ebegin + b NE + le;
    // tile_conn[(b NE + le) nen + a] is node a of that element.
    const int32_t *tile_elem;
    const int32_t *tile_conn;
};

// Fused assembly: the element kernel of each tile (a compact patch of NE
// elements of one connected component) sums, in its LDS output stage, the
// blocks of every upper node pair its elements touch, in ascending element
// order, and writes the pairs whose contributing elements all lie in the
// tile straight into the CSR values; the other (boundary) pairs leave a 6 x 6
// partial sum that pair_finalize_kernel adds up in tile order.
//   tp_ptr[t] .. tp_ptr[t+1]   the pair entries of tile t, ordered by (I, J)
//   trec[3q .. 3q+2]  the CSR placement of the entry's pair (I, J), so that
//                 its writes need no other table: per node {CSR position of
//                 the first free row, of the first fixed row (-1: none), row
//                 length, codes} for I then J, where codes holds 4 bits per
//                 comp c (0..5 offset in the node's free run, 8..13 in its
//                 fixed run, 15 not in the block) and bit 30 of I's word
//                 marks I = J; then {offset of J's free run in I's rows, of
//                 J's fixed run, of I's free run in J's rows, of I's fixed
//                 run} (-1: none)
//   tcontrib[q] = contribution count (bits 0..2) and up to six contributions
//                 of 10 bits from bit 3 on: tile slot le (5 bits) | local block
//                 (4 bits) << 5 | transposed << 9, in element order
//   tslot[q]    = partial slot of a boundary entry, -1 for a pair completed
//                 in the tile
//   brec[3b .. 3b+2], bp_ptr[b] .. bp_ptr[b+1]: the boundary pairs
//                 (placement records as trec) and their partial slots, which
//                 are consecutive per pair and in tile order
//   partial     nslots x 36 FP64 partial sums (row-major 6 x 6, rows of I)
struct TilePlan {
    int64_t ntiles;
    int64_t nbnd;                 // boundary pairs
    const int32_t *tp_ptr;
    const int4 *trec;
    const uint64_t *tcontrib;
    const int32_t *tslot;
    const int4 *brec;
    const int32_t *bp_ptr;
    double *partial;
};

// Mixed-precision operator of one model for the explicit step, on node-major
// dof positions R = 6 i + c (all dofs; fixed positions carry invMC = 0 and
// stay at zero).
//   kmb    membrane + bending stiffness, FP32, in slices of 10 nodes (60
//          rows, one wavefront per slice, lane = 6 (i mod 10) + c): slice g
//          holds nsl_g neighbour slots and the value pair (2 c2, 2 c2 + 1) of
//          slot s for lane l sits at kmb[goff[g] + (3 s + c2) * 60 + l]
//          (float2 units).  Slot s of a row refers to the s-th neighbour of
//          its node (nbr lists, sorted, including the node); slots past that
//          node's count are zero.
//   kml    residual of each kmb value, int16, at the same index (short2 per
//          float2): value = hi + lo 2^(e - 165) with e the biased exponent of
//          the FP32 value hi (zero for hi = 0), i.e. the FP64 value to a
//          relative 2^-39 in 6 bytes
//   cs     transverse-shear factor rows per element (FP64), K_s = sum C^T C
//   dr     drilling 3 x 3 rotation block per node, FP64, symmetric storage
//          (00, 01, 02, 11, 12, 22)
// Each value of K_mb is stored as FP32 plus an int16 residual (kml), which
// restores it to a relative 2^-39; the products run in FP64.
struct MixedOp {
    int64_t nrows;            // 6 N
    int32_t nnodes;
    int32_t nelem;
    const int64_t *goff;      // nslices + 1
    const float2 *kmb;
    const short2 *kml;        // residuals of kmb
    const int32_t *nbr_ptr;   // N + 1
    const int32_t *nbr_node;
    const double *dr;         // N x 6
    const int32_t *ne_ptr;    // node -> incident elements (ascending)
    const int32_t *ne_pack;   // per incidence: element, its nen nodes (stride 4 for T3, 8 for Q4)
    const double *cs;         // E x nsr x 6 nen
};

// Scatter plan of one block (ALL or FF) -- see scatter_pair_kernel.
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
    // upper pairs (I <= J): nodes and the neighbour slots kI (J in I's list)
    // and kJ (I in J's list), for the pair-driven scatter
    int64_t npairs;
    int64_t pbegin, pend;         // pair range of this launch, [pbegin, pend)
    const int32_t *pair_node;     // 2 * npairs: I, J
    const int32_t *pair_slot;     // 2 * npairs: kI, kJ
};

// Parameters of one stress-resultant launch (inspectintegpoints): one thread
// per (listed element, integration point).
//   g_mat  homogeneous: per group 16 doubles -- Dps 9 (row-major), 5/6 Dt 4,
//          thickness, stab_c, unused; layered: per group 41 -- A 9, B 9, D 9,
//          H 4 (row-major), layup csys 9 (row-major [i][j]), thickness
//   ocs    output csys (row-major [i][j]): one matrix (ocs_kind 1) or
//          nlist x npts matrices (ocs_kind 2); ocs_kind 0 takes E_G
//   out    nlist x npts x 8: N11 N22 N12 M11 M22 M12 Q1 Q2
//   status [0] singular metric, [1] output csys normal off the element normal
struct ResParams {
    int32_t nlist;
    int32_t nelem;
    int32_t ocs_kind;
    int32_t twice;          // layered Q4: displacements transformed twice (reference form)
    double stab_c;          // layered shells
    const int32_t *felist;  // nlist, or nullptr for elements 0..nlist-1
    const double *u;        // N x 6 nodal displacements and rotations (global basis)
    const double *ocs;
    const double *g_mat;
    const int32_t *elem_group;  // nelem or nullptr
    double *out;
    int *status;
};

// Rows of the free-free CSR matrix for the row products of the time loop and
// the power iteration.  Each row of a shell matrix is a sequence of runs of
// consecutive columns, one run per neighbour node (its free dofs), so the
// products need no per-entry column index:
//   rowrec[R]  = {first entry of row R, first and end neighbour slot of the
//                 row's node, 0}
//   slotrec[k] = {(offset of neighbour k's run within the row) << 4 | run
//                 length, first column of the run}
struct RunRows {
    int64_t n;
    const int4 *rowrec;
    const int2 *slotrec;
};

// Nodes of the free-free block for the node-grouped step (cd_node_step_kernel):
// one record per node with at least one free dof, in node order.
//   rec[i]  = {CSR position of the node's first row, that row (first free
//              dof d0), first and end neighbour slot}
//   info[i] = row length << 3 | number of rows (free dofs, rows d0 .. d0+nr-1)
struct NodeRows {
    int64_t n;
    const int4 *rec;
    const int32_t *info;
};

}  // namespace fs
