/*
 * flexshell.h -- C ABI of the MI355X shell finite-element library.
 *
 * The library computes, on one gfx950 GPU, the hot path of the shell FEMMs of
 * FinEtoolsFlexStructures.jl (v3.6.4): nodal normals, element stiffness of the
 * T3FF / T3FFComp / Q4RS / Q4RSComp shells scattered into a CSR global matrix,
 * the lumped mass, and the explicit central-difference loop built on them.
 *
 * Every entry point is a plain C function taking plain pointers and sizes; no
 * C++ or GPU types cross the boundary.  Host arrays are read during the call
 * and copied to the device; nothing is retained after return.  Every function
 * returns FS_OK (0) or a negative status; fs_last_error() then holds a
 * message.  Calls never unwind across the ABI.  Contexts are independent and
 * may be used from different threads; one context must not be used from two
 * threads at once.  All arrays use the memory order of the reference's Julia
 * data: N x 3 / N x 6 node arrays are column-major (x[], y[], z[] planes),
 * connectivity is E x nen row-major (Vector{NTuple{nen,Int}}).
 *
 * The reference interfaces each entry point replaces are cited per function
 * (paths relative to the FinEtoolsFlexStructures.jl repository root).
 */
#ifndef FLEXSHELL_H
#define FLEXSHELL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define FS_OK 0
#define FS_ERR_ARG (-1)         /* invalid argument (sizes, indices, NULL) */
#define FS_ERR_STATE (-2)       /* call sequence violated (e.g. no normals) */
#define FS_ERR_DEVICE (-3)      /* HIP runtime error */
#define FS_ERR_NOMEM (-4)       /* device or host allocation failed */
#define FS_ERR_NUMERIC (-5)     /* singular metric, non-positive material */
#define FS_ERR_UNSUPPORTED (-6) /* valid input outside what is implemented */

/* ---- element families ---- */
#define FS_T3 3 /* three-node flat facet: T3FF / T3FFComp */
#define FS_Q4 4 /* four-node quadrilateral: Q4RS / Q4RSComp */

/* ---- matrix blocks (which rows/columns are assembled) ---- */
#define FS_BLOCK_ALL 0 /* all dofs, nalldofs x nalldofs (SysmatAssemblerSparse*) */
#define FS_BLOCK_FF 1  /* free-free block, nfreedofs x nfreedofs (SysmatAssemblerFFBlock) */

/* ---- transverse shear formulations of the T3 family ---- */
#define FS_SHEAR_AVERAGE_B 0 /* __TRANSV_SHEAR_FORMULATION_AVERAGE_B (default) */
#define FS_SHEAR_AVERAGE_K 1 /* __TRANSV_SHEAR_FORMULATION_AVERAGE_K */

/* ---- explicit loop precision ---- */
#define FS_PREC_FP64 0  /* all products in FP64 */
#define FS_PREC_MIXED 1 /* membrane+bending stiffness held in FP32 (products
                           widened to and summed in FP64); state, mass, shear
                           and drilling stabilisation terms in FP64 */
#define FS_PREC_MIXED_F32 2 /* as FS_PREC_MIXED with the transverse-shear
                               stiffness also held in FP32 (in the membrane +
                               bending slices): fewer bytes per step; its
                               rounding grows with the element size over the
                               thickness, so it suits meshes with elements of
                               a few thicknesses or less */

typedef struct fs_ctx fs_ctx;
typedef struct fs_model fs_model;

/*
 * Material and formulation parameters of one stiffness evaluation.
 * Replaces the FEMM fields read at the start of stiffness():
 * FEMMShellT3FFModule.jl:657-663 (material -> Dps, Dt; drilling scale;
 * formulation; stab_fun), FEMMShellQ4RSModule.jl:901-907, and for layered
 * shells the per-group laminate data FEMMShellT3FFCompModule.jl:600-605,
 * FEMMShellQ4RSCompModule.jl:903-907.
 *
 * stab_fun(t, h) crosses the ABI as its coefficient c in t^2/(t^2 + c h^2):
 * the defaults are 5/12/1.5 (T3, FEMMShellT3FFModule.jl:36) and 0.1 (Q4,
 * FEMMShellQ4RSModule.jl:32).
 */
typedef struct fs_shell_params {
    int32_t composite;        /* 0: homogeneous (Dps, Dt, thickness); 1: layered */
    int32_t formulation;      /* FS_SHEAR_AVERAGE_B / _K (T3 only) */
    double thickness;         /* homogeneous: constant otherdimension */
    double stab_c;            /* Lyly stabilisation coefficient c */
    double drilling_stiffness_scale;
    double Dps[9];            /* homogeneous plane-stress 3x3, column-major */
    double Dt[4];             /* homogeneous transverse shear 2x2 (the 5/6 factor of
                                 FEMMShellT3FFModule.jl:658-659 is applied inside) */
    int32_t ngroups;          /* layered: number of layup groups */
    int32_t reserved;
    const double *group_ABDH;      /* ngroups x 31: A(9) B(9) D(9) H(4), column-major */
    const double *group_csys;      /* ngroups x 9: layup csys matrix (column-major) */
    const double *group_thickness; /* ngroups: thickness(layup) */
    const int32_t *elem_group;     /* nelem 0-based group ids, NULL = all group 0 */
    /* Homogeneous shells with several materials (a batch of independent
     * models): if not NULL, ngroups x 16 doubles per group -- Dps(9) Dt(4)
     * (column-major), thickness, stab_c, drilling_stiffness_scale -- replace
     * the single-material fields above for the elements of each group. */
    const double *group_homog;
} fs_shell_params;

/*
 * Lumped-mass parameters.  Replaces mass() FEMMShellT3FFModule.jl:772-791,
 * FEMMShellQ4RSModule.jl:981-1002 (rho, thickness) and the composite forms
 * FEMMShellT3FFCompModule.jl:737, FEMMShellQ4RSCompModule.jl:1002
 * (laminate_inertia! per group).
 */
typedef struct fs_mass_params {
    int32_t composite;
    int32_t ngroups;
    double rho;                  /* homogeneous mass density */
    double thickness;            /* homogeneous thickness */
    const double *group_inertia; /* ngroups x 2: (mass density, moment of inertia density) */
    const int32_t *elem_group;   /* nelem, NULL = all group 0 */
    const double *group_rho_t;   /* homogeneous groups: ngroups x 2 (rho, thickness), or NULL */
} fs_mass_params;

/*
 * Central-difference loop parameters.  Replaces _cd_loop!
 * (examples/shells/dynamics/homogeneous/explicit/plate_expl_examples.jl:61-94).
 * The force closure force!(F, t) crosses the ABI as F(t_n) = g[n] * Fmag with
 * the caller sampling g at t_n = t_{n-1} + dt, n = 0..nsteps.
 */
typedef struct fs_cd_params {
    double dt;          /* time step, after the tend/dt adjustment of :72-75 */
    double cdamp;       /* C = cdamp * diag(M)  (ksi * 2 * omegad); 0 = undamped */
    int64_t nsteps;
    int32_t precision;  /* FS_PREC_FP64, FS_PREC_MIXED or FS_PREC_MIXED_F32 */
    int32_t probe_stride; /* record probes every probe_stride steps (>= 1) */
    /* Batches of independent models stepped together: with ngroups > 0 each
     * free row r belongs to model row_group[r], which has its own time step
     * group_dt[g] and damping group_cdamp[g]; g then holds (nsteps+1) x
     * ngroups samples (g[n * ngroups + model]) and dt/cdamp above are unused. */
    int32_t ngroups;
    int32_t use_graph;  /* 1: capture the step loop in a hipGraph and launch it once */
    const int32_t *row_group;
    const double *group_dt;
    const double *group_cdamp;
    /* dt of invMC = 1 / (M + dt_mass/2 C): _cd_loop! forms invMC (:70-71)
     * from the dt it was called with, before the adjustment of :72-75, so a
     * damped loop whose dt was adjusted uses both.  0 = dt.  Batches use
     * group_dt for both. */
    double dt_mass;
} fs_cd_params;

/* ---------------------------------------------------------------------- */
/* Context                                                                 */
/* ---------------------------------------------------------------------- */

/* Create a context on HIP device `device` (one stream per context). */
int fs_ctx_create(int device, fs_ctx **out);
/* Release the caller's handle.  Models of the context keep it alive: the
 * context is torn down when the handle and every model are gone, so
 * contexts and models may be destroyed in any order (after fs_ctx_destroy
 * the only valid call on its models is fs_model_destroy).  Any call that
 * fails returns with the context's queued work finished. */
void fs_ctx_destroy(fs_ctx *ctx);
/* Message of the last failing call on this context ("" if none). */
const char *fs_last_error(const fs_ctx *ctx);
/* Library version string. */
const char *fs_version(void);

/* ---------------------------------------------------------------------- */
/* Model (mesh + dofs + device-resident matrices)                          */
/* ---------------------------------------------------------------------- */

/*
 * Upload a shell mesh.  Replaces the data gathered by gathervalues_asmat! /
 * fes.conn in the element loops (FEMMShellT3FFModule.jl:671,
 * FEMMShellQ4RSModule.jl:915).  xyz: nnodes x 3 column-major (geom0.values);
 * conn: nelem x nen row-major, node numbers with the given index_base (1 for
 * Julia).  nen = FS_T3 or FS_Q4.
 */
int fs_model_create(fs_ctx *ctx, int nen, int64_t nnodes, const double *xyz, int64_t nelem,
                    const int64_t *conn, int index_base, fs_model **out);
void fs_model_destroy(fs_model *m);

/*
 * Upload the equation numbers (NodalField dchi.dofnums after numberdofs!):
 * nnodes x 6 column-major, numbers with index_base, free dofs numbered
 * 1..nfreedofs and fixed dofs after them.  The dofs of one node must be
 * consecutive within the free and within the fixed numbers (as numberdofs!
 * produces, with or without a node permutation).
 */
int fs_model_set_dofs(fs_model *m, const int64_t *dofnums, int64_t nalldofs, int64_t nfreedofs,
                      int index_base);

/*
 * Isoparametric nodal normals on the device: associategeometry! of
 * FEMMShellT3FFModule.jl:570-616 (nen = 3) or FEMMShellQ4RSModule.jl:472-525
 * (nen = 4).  normals_out (nnodes x 3 column-major) and valid_out (nnodes)
 * may be NULL; the normals stay resident for the stiffness call.
 */
int fs_model_normals_isoparametric(fs_model *m, double threshold_angle_deg, double *normals_out,
                                   uint8_t *valid_out);

/* associategeometry! of the layered FEMMs (FEMMShellT3FFCompModule.jl:500-521,
 * FEMMShellQ4RSCompModule.jl:462-503), whose layup coordinate systems are
 * constant per layup group: the nodal normal is the normalised sum, over the
 * incident elements in element order, of the element group's csys normal
 * (weighted by the corner Jacobian for Q4); a node is invalid where that sum
 * departs from the element normal (T3) or from the csys normal (Q4) by more
 * than the threshold.  group_normal: ngroups x 3 row-major (third column of
 * each group's csys); elem_group: nelem group indices, or NULL (all group 0).
 * Outputs as for fs_model_normals_isoparametric. */
int fs_model_normals_layered(fs_model *m, double threshold_angle_deg, int32_t ngroups, const double *group_normal,
                             const int32_t *elem_group, double *normals_out, uint8_t *valid_out);

/*
 * Upload normals computed on the host: the mcsys / layup.csys paths of
 * associategeometry! (FEMMShellT3FFModule.jl:584 with a user CSys,
 * FEMMShellT3FFCompModule.jl:509, FEMMShellQ4RSCompModule.jl:471).
 */
int fs_model_normals_upload(fs_model *m, const double *normals, const uint8_t *valid);

/*
 * Build the CSR sparsity pattern and the element-to-CSR scatter plan for a
 * block.  The pattern is the structural union of all element dof pairs
 * (zeros retained), as FinEtools' COO assembly followed by
 * SparseMatricesCSR.sparsecsr produces it (AssemblyModule.jl:47-52).
 */
int fs_model_pattern(fs_model *m, int block, int64_t *nrows, int64_t *nnz);
/* Copy the pattern out: rowptr (nrows+1), colval (nnz), with index_base. */
int fs_model_pattern_get(fs_model *m, int block, int64_t *rowptr, int64_t *colval, int index_base);

/*
 * Stiffness: stiffness(femm, assembler, geom0, u1, Rfield1, dchi) of
 * FEMMShellT3FFModule.jl:635-736, FEMMShellT3FFCompModule.jl:561-689,
 * FEMMShellQ4RSModule.jl:877-947, FEMMShellQ4RSCompModule.jl:861-958 with the
 * assembly of AssemblyModule.jl:20-53 (startassembly!/assemble!/makematrix!).
 * Builds the block's pattern if needed, computes the CSR values on the
 * device (kept resident for fs_model_explicit) and, if nzval_out is not
 * NULL, copies them to the host (nnz doubles, in pattern order).
 */
int fs_model_stiffness(fs_model *m, const fs_shell_params *p, int block, double *nzval_out);

/*
 * Lumped mass: mass(femm, SysmatAssemblerSparseDiag or FFBlock, geom0, dchi)
 * of FEMMShellT3FFModule.jl:757-811 and the other three FEMMs.  Writes the
 * diagonal (nrows of the block) to diag_out if not NULL; kept resident.
 */
int fs_model_mass(fs_model *m, const fs_mass_params *p, int block, double *diag_out);

/*
 * Explicit central-difference loop on the resident FF-block stiffness and
 * mass (_cd_loop!, plate_expl_examples.jl:61-94; ThreadedSparseCSR.bmul! at
 * :87 is the CSR product).  U0, V0, Fmag: nfreedofs; g: nsteps+1 samples.
 * probe_out receives U[probe_dofs] (0-based) at steps 0, s, 2s, ... (s =
 * probe_stride): (nsteps/s + 1) x nprobes doubles.  U_out/V_out (may be NULL)
 * receive the final state.
 */
int fs_model_explicit(fs_model *m, const fs_cd_params *p, const double *U0, const double *V0,
                      const double *Fmag, const double *g, const int64_t *probe_dofs, int32_t nprobes,
                      double *probe_out, double *U_out, double *V_out);

/*
 * Largest generalized eigenvalue omega_max of (K_ff, M_ff) by power
 * iteration on the device, the form of pwr (examples/shells/dynamics/
 * homogeneous/explicit/clamp_cyl_transient_examples.jl:45-60) from the start
 * vector v0 (nfreedofs; the reference draws rand()).
 */
int fs_model_omega_max(fs_model *m, const double *v0, int32_t maxit, double *omega_max);

/*
 * fs_model_omega_max without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * arguments and result (bitwise); its work vectors live in a per-device
 * workspace that the library holds until the process ends (24 MiB + 4 KiB),
 * so the free-free block may have at most 1,048,576 rows (FS_ERR_UNSUPPORTED
 * beyond).  Every call on a device shares that workspace: calls that
 * overlap on one device, from any context, race.
 */
int fs_model_omega_max_noalloc(fs_model *m, const double *v0, int32_t maxit, double *omega_max);

/*
 * Element matrices before assembly: the elmat that stiffness() hands to
 * assemble!(assembler, elmat, dofnums, dofnums) (FEMMShellT3FFModule.jl:733,
 * FEMMShellQ4RSModule.jl:944), for the generic-assembler path and for tests.
 * blocks_out: nelem x nblk x 36 doubles, nblk = nen(nen+1)/2 upper node-pair
 * blocks (a, b), a <= b, in the order (0,0),(0,1),..,(1,1),..; each block is
 * 6 x 6 row-major (rows: dofs of local node a, columns: dofs of node b).
 */
int fs_model_element_blocks(fs_model *m, const fs_shell_params *p, double *blocks_out);

/*
 * numberdofs! on the device: free dofs first, then the fixed ones, nodes in
 * their order or in the order of a permutation (FinEtools NodalField
 * numberdofs!(dchi) and numberdofs!(dchi, perm) after setebc!/applyebc!:
 * test_shell_statics.jl:205, plate_expl_examples.jl:147), and the equation
 * numbers installed as fs_model_set_dofs would.
 *   is_fixed     N x 6 flags (row-major: node, dof), host
 *   perm         N 0-based node numbers in numbering order, or NULL
 *   dofnums_out  N x 6 0-based numbers, column-major like fs_model_set_dofs'
 *                input (may be NULL)
 *   nfree_out    number of free dofs (may be NULL)
 */
int fs_model_number_dofs(fs_model *m, const uint8_t *is_fixed, const int64_t *perm, int64_t *dofnums_out,
                         int64_t *nfree_out);

/* fs_model_surface_load kinds */
#define FS_LOAD_UNIFORM 0   /* one 6-vector force intensity for the whole mesh */
#define FS_LOAD_PER_POINT 1 /* E x npts x 6 values at the rule's points (fs_model_load_points) */

/*
 * Consistent nodal loads of a distributed surface force: distribloads(
 * FEMMBase(IntegDomain(fes, rule)), geom0, dchi, ForceIntensity(f), 2)
 * (test_shell_statics.jl:84-86; plate_expl_examples.jl:197-199).  npts picks
 * the rule: 1, 3 or 6 for TriRule(npts) on T3, 4 for GaussRule(2, 2) on Q4.
 * Each element's vector sum_q N f Jac w is summed into the nodes in element
 * order, as the element-by-element assembly does.  block FS_BLOCK_FF gives
 * the free rows (SysvecAssemblerFBlock, F[1:nfreedofs]), FS_BLOCK_ALL all
 * rows.  F_out: host.
 */
int fs_model_surface_load(fs_model *m, int32_t npts, int32_t kind, const double *force, int block, double *F_out);

/* Locations of the integration points of fs_model_surface_load (E x npts x 3,
 * host), where a position-dependent ForceIntensity is evaluated
 * (updateforce!(fi, loc, J, label, qp), e.g. computetrac! of
 * plate_expl_examples.jl:184-194); distribloads computes them internally. */
int fs_model_load_points(fs_model *m, int32_t npts, double *xyz_out);

/* Output coordinate systems of fs_model_resultants. */
#define FS_OCSYS_ISOPARAMETRIC 0 /* E_G of each point: the default mcsys of the
                                    homogeneous FEMMs (isoparametric!, and
                                    _isoparametric! FEMMShellQ4RSModule.jl:201-203) */
#define FS_OCSYS_MATRIX 1        /* one 3x3 matrix (column-major) for every point:
                                    a constant CSys, e.g. the default output csys
                                    of the layered FEMMs, layup_groups[1][1].csys */
#define FS_OCSYS_PER_POINT 2     /* nlist x npts column-major matrices that the caller
                                    evaluated with updatecsmat!(outputcsys, ...) */
/* fs_model_resultants flags */
#define FS_RES_LAYERED_Q4_SINGLE_T 1 /* Q4RSComp: transform the displacements into the
                                        element frame once (see below) */

/*
 * Stress resultants at the integration points: inspectintegpoints for the
 * three quantities :membrane_force, :bending_moment and :transverse_shear at
 * once (FEMMShellT3FFModule.jl:850-962, FEMMShellQ4RSModule.jl:1061-1170,
 * FEMMShellT3FFCompModule.jl:809-943, FEMMShellQ4RSCompModule.jl:1077-1209).
 * The inspector callback stays with the caller, which hands it, for listed
 * element l and point j, out[l][j][0..2] (membrane), [3..5] (bending) or
 * [6..7] (shear).
 *   p         the material of stiffness(); formulation and drilling scale unused
 *   u         N x 6 nodal displacements and rotations (u.values, global basis,
 *             row-major), host memory
 *   felist    nlist 0-based element numbers (host); NULL = elements 0..nlist-1
 *   ocs_kind  FS_OCSYS_*, with ocs the matrices it needs (host)
 *   flags     0 reproduces the reference.  Its layered Q4 form multiplies the
 *             displacements already taken into the element frame,
 *             edisp_e = T u (FEMMShellQ4RSCompModule.jl:1163), by B T again
 *             (:1168-1170, :1196); FS_RES_LAYERED_Q4_SINGLE_T applies T once,
 *             as the other three FEMMs do
 *   out       nlist x npts x 8 (npts 1 for T3 at the centroid, 4 for Q4 at the
 *             Gauss points): N11 N22 N12 M11 M22 M12 Q1 Q2 in the output csys
 *   mismatch  optional: 1 if some output csys normal is less than 0.95 aligned
 *             with the element normal (the reference's "Coordinate systems
 *             mismatched?" warning), else 0
 * Requires the normals (associategeometry!).
 */
int fs_model_resultants(fs_model *m, const fs_shell_params *p, const double *u, int64_t nlist,
                        const int32_t *felist, int32_t ocs_kind, const double *ocs, int32_t flags, double *out,
                        int32_t *mismatch);

/* ---------------------------------------------------------------------- */
/* Generic assembler: element matrices computed on the host               */
/* ---------------------------------------------------------------------- */

typedef struct fs_coo fs_coo;

/*
 * The assembler protocol for a FEMM with no element kernel in this library
 * (FinEtools' SysmatAssemblerSparse, extended to CSR by
 * SysmatAssemblerSparseCSRSymm, AssemblyModule.jl:28-52): the caller buffers
 * what assemble!(a, mat, dofnums_row, dofnums_col) receives for each of the
 * nmat element matrices and hands the whole buffer over in makematrix!.
 * fs_coo_create takes the dof numbers (startassembly! sizes: nrows x ncols,
 * kr x kc per element matrix; dofnums_row nmat x kr and dofnums_col nmat x kc,
 * row-major, with index_base; numbers beyond nrows / ncols are dropped, as
 * SysmatAssemblerFFBlock drops the fixed dofs) and builds the pattern: the
 * structural union of the entries, zeros retained, columns sorted within
 * each row.  fs_coo_values reduces the element matrices (nmat x kr x kc, each
 * column-major as Julia stores it) on the device, summing the duplicates of
 * each entry in the order assemble! appended them (element order, column by
 * column), as sparse(I, J, V) does; nzval receives the nnz values (NULL:
 * they stay on the device).  A handle may be reused for new values with the
 * same dof numbers.
 */
int fs_coo_create(fs_ctx *ctx, int64_t nrows, int64_t ncols, int32_t kr, int32_t kc, int64_t nmat,
                  const int64_t *dofnums_row, const int64_t *dofnums_col, int index_base, fs_coo **out);
int fs_coo_pattern(const fs_coo *h, int64_t *nnz);
/* rowptr (nrows + 1) and colval (nnz) with index_base. */
int fs_coo_pattern_get(const fs_coo *h, int64_t *rowptr, int64_t *colval, int index_base);
int fs_coo_values(fs_coo *h, const double *elmats, double *nzval);
void fs_coo_destroy(fs_coo *h);

/* Device time (ms) of the last fs_model_stiffness / fs_model_explicit call,
 * measured with HIP events on the context stream: which = 0 element kernels
 * (first launch to last element range done; with the assembly pipeline this
 * includes the scatter work that overlaps them), 1 the rest of the scatter
 * (the whole scatter kernel when the pipeline has one range), 2 whole
 * stiffness (0 + 1), 3 explicit loop, 4 construction of the mixed-precision
 * operator (0 when the last loop ran in FP64 or reused an operator built
 * earlier), 5 the stress-resultant kernel of the last fs_model_resultants,
 * 6 host time of the last fs_model_explicit with use_graph spent capturing
 * and instantiating its step graph (0 when it replayed the model's cached
 * graph: same sizes, scalars, load factors and buffers as the graph's). */
int fs_model_last_timing(const fs_model *m, int which, double *ms);

/* Element-kernel launch shape chosen by the autotuner for a family (FS_T3 /
 * FS_Q4) and shear formulation on this context (no counterpart in the
 * reference, which has no GPU launch shapes): shape[0] = variant index (-1
 * before the first stiffness evaluation of that kind), shape[1] = elements
 * per workgroup, shape[2] = workgroup size.  The autotuner times every shape
 * once on the first evaluation and keeps the fastest; every shape computes
 * bitwise the same element blocks. */
int fs_ctx_tuning(const fs_ctx *ctx, int nen, int formulation, int32_t *shape);

/* Fix the launch shape (variant >= 0), re-arm the autotuner (-1), or switch
 * it off so that the default shape is used (-2). */
int fs_ctx_set_tuning(fs_ctx *ctx, int nen, int formulation, int32_t variant);

/* Element ranges of the assembly pipeline of fs_model_stiffness (no
 * counterpart in the reference, which assembles on the CPU): the element
 * kernel runs over `ranges` consecutive element ranges, and the scatter of
 * the node pairs completed by each range runs on a second stream while the
 * next range computes.  1 = element kernel, then scatter, on one stream.
 * Every setting gives bitwise the same matrix.  Default 1; 1 <= ranges <= 64. */
int fs_ctx_set_pipeline(fs_ctx *ctx, int ranges);

/* Form of the stiffness assembly of fs_model_stiffness (no counterpart in the
 * reference, which assembles element by element on the CPU):
 *   1 (default) fused element tiles: the elements are grouped into compact
 *     tiles (per connected component, along a Morton curve of their
 *     centroids); one kernel computes a tile's element matrices and writes
 *     the CSR values of every node pair whose contributing elements all lie
 *     in the tile, summed in element order; the other pairs leave partial
 *     sums that a second kernel adds in tile order.  No element matrices go
 *     through HBM.
 *   0 element node-pair blocks to HBM, then the pair scatter (sums in
 *     element order; the pipeline of fs_ctx_set_pipeline applies).
 * Both give the same pattern; the values agree to rounding (the boundary
 * pairs of form 1 add their tile partials).  Each form is bitwise
 * reproducible and gives a model bitwise the same matrix alone and in a
 * batch. */
int fs_ctx_set_assembly(fs_ctx *ctx, int mode);

/* Form of the row products of the FP64 explicit step and of the power
 * iteration of fs_model_omega_max (no counterpart in the reference, whose
 * products are CSR products on the host): 0 = four lanes per
 * row over the neighbour runs; 8 or 16 = that many lanes per node over the
 * node's rows together (consecutive lanes read consecutive values), each
 * lane loading positions_per_lane values of a row at once (8 lanes: 4..8,
 * 16 lanes: 2..4; 0 = chosen from the block's longest row).  All forms give
 * the same sums up to their order of addition; every setting is bitwise
 * reproducible and gives a model bitwise the same result alone and in a
 * batch.  Default 8 lanes per node, positions chosen. */
int fs_ctx_set_cd_lanes(fs_ctx *ctx, int lanes_per_node, int positions_per_lane);

/* Storage of the mixed-precision operator built by the last FS_PREC_MIXED /
 * FS_PREC_MIXED_F32 loop (no counterpart in the reference, whose loop runs on
 * the FP64 CSR matrix only): info[0] FP32 value pairs of the slices
 * (membrane + bending, plus shear for FS_PREC_MIXED_F32), info[1] FP64 shear
 * factor rows per element (0 for FS_PREC_MIXED_F32), info[2] entries of the
 * node neighbour lists, info[3] 64-row slices.  Zeros before the first mixed
 * loop. */
int fs_model_mixed_info(const fs_model *m, int64_t *info);

#ifdef __cplusplus
}
#endif

#endif /* FLEXSHELL_H */
