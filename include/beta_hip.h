/*
 * beta_hip.h -- C ABI of libbeta_hip.so, the MI355X (gfx950) implementation
 * of BETA's finite-element hot path: element integration, global assembly
 * into a symmetric sparse matrix, and the Jacobi-PCG solve.
 *
 * Every entry point cites the reference interface it replaces (paths are
 * relative to the reference root).  The reference's solver/storage plug-in
 * is LargeMatrix (beta/math/equation/largeMatrix.cpp, sparse.cpp), its
 * assembler plug-in is BasicAssembler (beta/assemblers/*), and the linear
 * system facade is Equations (beta/math/equation/equation.cpp).
 *
 * Conventions
 *  - The caller owns every host array; the context owns device buffers
 *    (mesh, pattern, matrix values, preconditioner, solver vectors).
 *  - Every call returns BHIP_OK (0) or a negative status; the message of the
 *    last failure is available from bhip_last_error().  The reference
 *    terminates the process in these cases (exit(1)/FatalError); the C++
 *    adapters in betanoox_amd/host reproduce that when asked to.
 *  - A context is used from one host thread at a time.
 *  - DOF numbering is the reference's: node-major, cumulative first DOF per
 *    node in node order (beta/mesh/node.cpp:206-214), element-local DOF order
 *    node-major (beta/elements/BasicElement.cpp:49-68).
 */
#ifndef BETA_HIP_H
#define BETA_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define BHIP_OK                 0
#define BHIP_ERR_ARG           -1   /* bad argument or shape */
#define BHIP_ERR_HIP           -2   /* HIP runtime failure */
#define BHIP_ERR_OOM           -3   /* device allocation failed */
#define BHIP_ERR_NOT_CONVERGED -4   /* sparse.cpp:225-229 */
#define BHIP_ERR_ZERO_DIAG     -5   /* largeMatrix.cpp:55-62 */
#define BHIP_ERR_NEG_DIAG      -6   /* largeMatrix.cpp:50-53, BasicElement.cpp:116-130 */
#define BHIP_ERR_NEG_DETJ      -7   /* ElasticityElement3D.cpp:653-670 */
#define BHIP_ERR_STATE         -8   /* call out of order (e.g. solve before assemble) */
#define BHIP_ERR_UNSUPPORTED   -9   /* element/material combination not built */
#define BHIP_ERR_CAPACITY     -10   /* size beyond a documented limit */

/* ------------------------------------------------------------- materials */
#define BHIP_MAT_ISO_ELASTIC    1   /* 3D: E, nu                      (elastic.cpp:335-376, isotropic input) */
#define BHIP_MAT_CMAT3D         2   /* 3D: 36 doubles, row-major C    (ElasticMaterial::globalCmat) */
#define BHIP_MAT_PLANE_STRESS   3   /* 2D: e11 e22 nu12 G12           (pStress.cpp:49-117) */
#define BHIP_MAT_HEAT           4   /* 3D scalar: k11 k22 k33 capacity (heatTransferMaterial.cpp:65-136) */
#define BHIP_MAT_J2             5   /* 3D: E nu sigma0 H              (plastic.cpp, POWER curve m_n=1) */

/* -------------------------------------------------------------- assembly */
#define BHIP_ASM_STIFFNESS      1   /* update("K") + addMatrix(Ke, map, size, NULL) */
#define BHIP_ASM_HEAT_TRANSIENT 2   /* update("KM"), formK_s, formF_s_new, addMatrix(Mbar, map, size, Fbar) */
#define BHIP_ASM_PLASTIC_TANGENT 3  /* update("KI") with the consistent tangent from the last return map */

/* ----------------------------------------------------------------- solve */
#define BHIP_STOP_REF_R_OVER_X  0   /* sqrt(|r|^2/|x|^2) < tol   (sparse.cpp:223) */
#define BHIP_STOP_REL_RESID     1   /* |b - K x| / |b| <= tol     (true residual checked at exit) */

#define BHIP_PREC_FP64          0   /* plain FP64 Jacobi-PCG on FP64 CSR */
#define BHIP_PREC_FP32_IR       1   /* FP32 inner Jacobi-PCG (FP32 values), FP64 iterative refinement */
#define BHIP_PREC_FP64_BLOCKED  2   /* plain FP64 Jacobi-PCG on FP64 B x B node blocks (SELL-64 layout of
                                       the FP32 inner solve); needs the same 1-3 DOFs at every node */

typedef struct bhip_ctx bhip_ctx;

typedef struct bhip_solve_opts {
    double tol;          /* stopping tolerance (reference default 1e-4, equation.cpp:1566-1567) */
    int    maxit;        /* iteration cap (reference default 10000) */
    int    stop_rule;    /* BHIP_STOP_* */
    int    precision;    /* BHIP_PREC_* */
    double inner_tol;    /* FP32_IR: relative residual reduction per inner solve (loosened, up to
                            0.1, for a step whose outer residual is already near tol) */
    int    max_outer;    /* FP32_IR: refinement cap */
    int    check_every;  /* host polls convergence every this many iterations (0 = 32) */
    int    use_x0;       /* 1: start from the x0 held by the context (bhip_set_x0), 0: from zero */
} bhip_solve_opts;

typedef struct bhip_stats {
    int    iters;            /* PCG iterations (inner iterations for FP32_IR) */
    int    outer_iters;      /* refinement steps (FP32_IR), 1 for FP64 */
    double ref_norm;         /* final sqrt(|r|^2/|x|^2) of the recursive residual */
    double rel_resid;        /* final |b - K x|_2 / |b|_2, true residual in FP64 */
    double ms_total;         /* wall time of the solve on the device (HIP events) */
    double ms_spmv_avg;      /* mean duration of one SpMV kernel inside the solve */
    int    spmv_launches;    /* SpMV kernels launched in the timed region */
    double spmv_bytes;       /* algorithmic bytes of one SpMV */
    long long case_iters;    /* iterations summed over load cases (bhip_solve_cases: a packed
                                solve advances all its cases per iteration, so iters counts
                                the longest case and case_iters every case); = iters otherwise */
} bhip_stats;

/* --------------------------------------------------------------- context */
/* LargeMatrix ctor (largeMatrix.cpp:13-24) + Equations ctor (equation.cpp:1540-1570). */
int         bhip_create(int device, bhip_ctx **out);
void        bhip_destroy(bhip_ctx *ctx);
const char *bhip_last_error(const bhip_ctx *ctx);
const char *bhip_version(void);

/* ------------------------------------------------------------------ mesh */
/* BasicMesh::ReadMesh (BasicMesh.cpp:79-190) data + Node::setNumDofPerNode
 * (node.cpp:158-227).  xyz: n_nodes*dim doubles; conn: n_el*nnpe ints in the
 * reference node order (hex8: BasicElement.cpp:518-525; quad4 CCW);
 * ndof_per_node NULL means every node has ndof_uniform DOFs.  Per-node counts
 * may differ only on nodes outside every element (the stiffness assembly and
 * the scalar CSR solves take such a mesh; the blocked forms, load cases,
 * internal forces, heat and plasticity need the same count at every node). */
int bhip_set_mesh(bhip_ctx *ctx, int dim, int n_nodes, const double *xyz, int n_el, int nnpe,
                  const int *conn, const int *ndof_per_node, int ndof_uniform);
int bhip_num_dof(const bhip_ctx *ctx);
/* first DOF of every node (Node::getFirstDof) */
int bhip_get_dof_first(bhip_ctx *ctx, int *dof_first);

/* ----------------------------------------------------------- restraints */
/* Equations::setRestraint + setSolution (equation.cpp:306-341): restrained[i]
 * != 0 marks DOF i, values[i] is its prescribed value.  Changing the set of
 * restrained DOFs invalidates the pattern; bhip_set_prescribed only changes
 * the values. */
int bhip_set_restraints(bhip_ctx *ctx, const unsigned char *restrained, const double *values);
int bhip_set_prescribed(bhip_ctx *ctx, const double *values);

/* -------------------------------------------------------------- pattern */
/* BasicModel::allocateEquationsMatrix (BasicModel.cpp:1022-1033) ->
 * Equations::specifyNonZeroLocation (equation.cpp:1216-1267) ->
 * SymmetricSparseMatrix::specifyNonZeroLocation/allocate (sparse.cpp:479-495),
 * built on the device once. */
int bhip_build_pattern(bhip_ctx *ctx);
int bhip_pattern_nnz(const bhip_ctx *ctx, long long *nnz_lower, long long *nnz_full);
/* The reference storage layout: lower triangle, sorted columns, diagonal last
 * in every row (sparse.cpp:144), rowptr has n_dof+1 entries. */
int bhip_get_pattern_lower(bhip_ctx *ctx, int *rowptr, int *colind);
int bhip_get_values_lower(bhip_ctx *ctx, double *val);
/* Full symmetric CSR as stored on the device (int32 indices). */
int bhip_get_csr(bhip_ctx *ctx, int *rowptr, int *colind, double *val);

/* ------------------------------------------------------------ materials */
/* MaterialFactory + SetElementProperty (Factory_Material.cpp:13-49,
 * BasicModel.cpp:664-771): params holds n_groups records of the kind's
 * length; elem_group[e] selects the record of element e (NULL: element e uses
 * record e when n_groups == n_el, record 0 when n_groups == 1). */
int bhip_set_materials(bhip_ctx *ctx, int kind, int n_groups, const double *params,
                       const int *elem_group);
/* IsoElement::setIntegrationOrder (isoElement.cpp:463-478) and
 * getIntegrationFactor() (thickness in 2D, 1 in 3D). */
int bhip_set_integration(bhip_ctx *ctx, int order, double factor);

/* ------------------------------------------------------------- assembly */
/* SerialAssembler::assembleK (SerialAssembler.cpp:35-59) with
 * Equations::addMatrix (equation.cpp:445-562): zeroes K and the pseudo-load,
 * integrates every element and scatters it with restrained DOFs eliminated
 * (identity rows, lifted columns).  coef: BHIP_ASM_HEAT_TRANSIENT takes
 * {alpha, dt} and u_state = u_s (n_dof); others pass NULL. */
int bhip_assemble(bhip_ctx *ctx, int what, const double *coef, const double *u_state);
/* Equations::pseudoLoadVector after assembly (lift + element forces). */
int bhip_get_pseudo_load(bhip_ctx *ctx, double *f);
/* Equations::loadVector (addLoadToEqn, equation.cpp:851-868). */
int bhip_set_load(bhip_ctx *ctx, const double *load);
/* Equations::solve RHS build (equation.cpp:1324-1342): b[i] = load+pseudo for
 * free i, the prescribed value for restrained i. */
int bhip_build_rhs(bhip_ctx *ctx);
int bhip_get_rhs(bhip_ctx *ctx, double *b);
int bhip_set_rhs(bhip_ctx *ctx, const double *b);
int bhip_set_x0(bhip_ctx *ctx, const double *x0);

/* ---------------------------------------------------------------- solve */
/* SymmetricSparseMatrix::solve (sparse.cpp:114-242) on the resident system:
 * rhs and x0 already on the device, solution left on the device. */
int bhip_solve(bhip_ctx *ctx, const bhip_solve_opts *opts, bhip_stats *stats);
int bhip_get_solution(bhip_ctx *ctx, double *x);
/* LargeMatrix::solve(double *b, double *x0): b holds the RHS on entry and the
 * solution on return; x0 may be NULL (zero initial guess). */
int bhip_pcg_solve(bhip_ctx *ctx, double *b, const double *x0, const bhip_solve_opts *opts,
                   bhip_stats *stats);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc): bhip_pcg_solve for
 * opts->precision == BHIP_PREC_FP64 only (other precisions: UNSUPPORTED), with
 * the same kernels, iterates and bits, launched without graphs.  The solver
 * state, diagonal flags and reduction records live in a fixed 4 MiB device
 * array per GPU (held until the process ends) instead of context buffers, so
 * the call allocates nothing once the context is assembled; a system needing
 * more than 524,224 reduction records (about 8.3M DOF at 81 entries per row)
 * is refused with BHIP_ERR_CAPACITY.  Zero diagonals are replaced as in
 * bhip_pcg_solve when a replacement is set.  Calls that overlap on one GPU
 * (other threads or contexts) share that array and race. */
int bhip_pcg_solve_noalloc(bhip_ctx *ctx, double *b, const double *x0, const bhip_solve_opts *opts,
                           bhip_stats *stats);
/* LargeMatrix::vectorProduct (sparse.cpp:452-478): y = K x. */
int bhip_spmv(bhip_ctx *ctx, const double *x, double *y);
/* BasicModel::getGlobalForces (BasicModel.cpp:1063-1110) for hex8
 * elasticity (BHIP_MAT_ISO_ELASTIC or BHIP_MAT_CMAT3D): for each of nvec
 * displacement vectors u (nvec x n_dof, full displacements with the
 * prescribed values on restrained DOFs), the internal force
 * force = sum_e f_e(u) over every DOF (ElasticityElement3D update("F"); on a
 * restrained DOF it is the reaction), optionally residual = load - force on
 * free DOFs and 0 on restrained ones (calculateResidualNorm,
 * equation.cpp:1516-1529) and max_abs = max |residual| per vector.  Applied
 * to the displacements of bhip_solve_cases it gives the reactions that
 * MultiFieldMacroElement::solveForStiffnessMatrixColumn takes as a column of
 * the macro-element stiffness (MultiFieldMacroElement.cpp:217-219).
 * residual and max_abs may be NULL. */
int bhip_internal_force(bhip_ctx *ctx, int nvec, const double *u, double *force, double *residual,
                        double *max_abs);
/* LargeMatrix::setReplaceZeroDiagonal_Option (largeMatrix.cpp:43-79, set by
 * the ReplaceZeroDiagonal command, equation.cpp:1819-1821): before a solve,
 * zero diagonal entries of the matrix are replaced by value (0, the default:
 * a zero diagonal is reported as BHIP_ERR_ZERO_DIAG).  As the reference, the
 * diagonal is scanned in row order: a negative entry before the first zero is
 * BHIP_ERR_NEG_DIAG. */
int bhip_set_zero_diag_replacement(bhip_ctx *ctx, double value);
/* Several load cases on the assembled stiffness: the multi-column solve of
 * MultiFieldMacroElement::CalculateK / solveForStiffnessMatrixColumn
 * (MultiFieldMacroElement.cpp:135-407), where each case re-lifts its prescribed
 * displacements through the assembled K (Equations::updatePseudoLoadVectorUsingGK,
 * equation.cpp:343-381) and is solved with the same matrix.
 * presc: nrhs x n_dof prescribed values (case-major; read on restrained DOFs);
 * load: n_dof external load shared by all cases, or NULL; x: nrhs x n_dof out.
 * 1 <= nrhs <= 8.  Needs a BHIP_ASM_STIFFNESS assembly.  With BHIP_PREC_FP32_IR
 * on a 3-DOF-per-node mesh the cases share one fused inner solve (each keeps its
 * own CG recurrence and stopping test); otherwise they are solved in turn.
 * stats: iters = inner iterations of the fused solve, rel_resid = worst case.
 * The context's own prescribed values, load and RHS are left unchanged. */
int bhip_solve_cases(bhip_ctx *ctx, int nrhs, const double *presc, const double *load, double *x,
                     const bhip_solve_opts *opts, bhip_stats *stats);
/* Packed solve: bhip_solve_cases for n assembled contexts at once.  The
 * reference solves one model's cases at a time (the CalculateK column loop,
 * MultiFieldMacroElement.cpp:135-407); here the FP32 inner PCGs of all n
 * contexts' cases advance together in one persistent set of kernels
 * (graph-captured chunks), each case keeping its own recurrence and stopping
 * test, every context to its own tolerance.  A context's results are the
 * same bits as bhip_solve_cases gives it alone.  Every context needs
 * BHIP_PREC_FP32_IR, 3 DOFs per node and a BHIP_ASM_STIFFNESS assembly, and
 * all share nrhs (1..8).  presc[i], load[i] (load or load[i] may be NULL) and
 * x[i] are laid out as in bhip_solve_cases; status[i] and stats[i] (stats may
 * be NULL) per context; returns the first failing status or BHIP_OK.  The
 * contexts must not be used by another thread meanwhile. */
typedef struct bhip_pack bhip_pack;
int  bhip_pack_create(int device, bhip_pack **out);
void bhip_pack_destroy(bhip_pack *pk);
int  bhip_pack_solve_cases(bhip_pack *pk, int n, bhip_ctx *const *ctxs, int nrhs, const double *const *presc,
                           const double *const *load, double *const *x, const bhip_solve_opts *opts,
                           bhip_stats *stats, int *status);
/* Time `reps` SpMVs on resident vectors: fp32 = 0 FP64 CSR, 1 the FP32 form of
 * the inner solve, 2 FP64 SELL blocks (BHIP_PREC_FP64_BLOCKED); returns the
 * mean kernel time in ms (HIP events) and the algorithmic bytes per SpMV. */
int bhip_bench_spmv(bhip_ctx *ctx, int fp32, int reps, double *ms_avg, double *bytes);
/* Tile shapes of the FP32 blocked (SELL-64) SpMV of the inner solve.  No
 * counterpart in the reference (its SpMV is the serial vectorProduct,
 * sparse.cpp:452-478).  A shape changes speed only: every shape sums each row
 * and the p.Ap partials in the same order, so solves give the same bits
 * whichever shape runs. */
typedef struct bhip_spmv_shape {
    int    id;               /* shape id (0: the default shape) */
    int    slices_per_block; /* 64-node slices (one wave each) per workgroup: 4, 8 or 16 */
    int    unroll;           /* block columns per loop trip; 0: runs of up to three union slots
                                per trip, x gathered once per run (matrices with three DOFs
                                per node whose layout has union slots; unroll 4 otherwise) */
    int    nontemporal;      /* 1: streaming loads of block values and columns */
    int    column_bytes;     /* bytes per stored block column of the tuned matrix: 2 (16-bit
                                offsets from the row's node) or 4 (node indices); 0 from
                                bhip_spmv_shape_describe */
    double ms;               /* best mean SpMV kernel time of the tuning rounds, timed with HIP
                                events around each launch inside inner PCG iterations */
    double gbps;             /* algorithmic bytes / ms */
    double iter_ms;          /* best mean time of one inner PCG iteration (SpMV + fused vector
                                kernels) with this shape; the tuner keeps the smallest */
} bhip_spmv_shape;
/* Autotuner: times every shape on the context's assembled matrix (three
 * interleaved rounds of `reps` inner PCG iterations per shape, so that the
 * caches hold what a solve leaves in them; the iterations leave the context's
 * solution untouched), keeps the shape with the fastest iteration for the
 * context's later solves and reports the table (up to max_entries rows).
 * n_entries = 0 and chosen = -1 when the matrix has no blocked FP32 form
 * (DOFs per node not uniform or above 3). */
int bhip_tune_spmv(bhip_ctx *ctx, int reps, bhip_spmv_shape *table, int max_entries, int *n_entries,
                   int *chosen);
/* Select a shape by id (a stored tuning result); BHIP_ERR_ARG for an unknown id. */
int bhip_set_spmv_shape(bhip_ctx *ctx, int id);
int bhip_get_spmv_shape(const bhip_ctx *ctx);
/* Shape parameters of an id (ms and gbps zero); BHIP_ERR_ARG past the last id. */
int bhip_spmv_shape_describe(int id, bhip_spmv_shape *info);
/* Diagnostic without a reference counterpart: a plain device copy of `bytes`
 * bytes (read + write), timed over `reps` launches; reports the achieved
 * read+write GB/s (STREAM-copy convention) of this GPU. */
int bhip_bench_stream(bhip_ctx *ctx, long long bytes, int reps, double *ms_avg, double *gbps);
/* Diagnostic without a reference counterpart: a read-only stream of `bytes`
 * bytes (16-byte nontemporal loads), timed over `reps` launches; reports the
 * read GB/s this GPU sustains, the practical ceiling of the read-dominated
 * SpMVs. */
int bhip_bench_read(bhip_ctx *ctx, long long bytes, int reps, double *ms_avg, double *gbps);

/* ------------------------------------------------------------ plasticity */
typedef struct bhip_plastic_stats {
    int    newton_iters;     /* linear solves in the increment */
    int    pcg_iters;        /* PCG iterations over those solves */
    double max_residual;     /* final max |load - resultant| over free DOFs */
    int    plastic_points;   /* Gauss points in the plastic region */
    int    max_return_iters; /* largest return-map iteration count */
    double ms;               /* device time of the increment */
} bhip_plastic_stats;

/* GaussISV allocation (createISVSpace, PlasticityModel.cpp:121-129): zero
 * stresses and effective plastic strain at every Gauss point; needs a J2
 * material, hex8 and the pattern. */
int bhip_plastic_init(bhip_ctx *ctx);
/* Tangent of BHIP_ASM_PLASTIC_TANGENT: the reference's continuum C_ep
 * (getTangentStiffness, plastic.cpp:583-610; the default) or the symmetric
 * part of the exact derivative of the return map as executed (no reference
 * counterpart: it leaves the stress update and so the converged state alone
 * and changes only how fast Newton gets there).  bhip_plastic_increment takes
 * the continuum tangent on the first pass of every increment either way;
 * bhip_plastic_update applies the selected one. */
#define BHIP_TANGENT_CONTINUUM   0
#define BHIP_TANGENT_ALGORITHMIC 1
int bhip_plastic_set_tangent(bhip_ctx *ctx, int tangent);
/* update("V") on every element (plasticElement.cpp:177-180): strain
 * increment from du (n_dof, increment displacement) and the FP64 return map
 * (PlasticMaterial::getStressIncrement, plastic.cpp:368-465). */
int bhip_plastic_update(bhip_ctx *ctx, const double *du);
/* getGlobalForces + calculateResidualNorm (BasicModel.cpp:1063-1110,
 * equation.cpp:1516-1529): residual = load - resultant of the stress
 * increments on free DOFs, 0 on restrained DOFs; max |residual|. */
int bhip_plastic_residual(bhip_ctx *ctx, double *residual, double *max_abs);
/* update("U") (GaussISV::updateISV, plastic.cpp:769-785) and accumulation of
 * the increment displacement into the total. */
int bhip_plastic_commit(bhip_ctx *ctx);
/* One load increment of PlasticityModel::DoPlasticAnalysis
 * (PlasticityModel.cpp:213-366) on the resident system: prescribed values
 * (n_dof, increment; NULL keeps the current ones), the applied load increment
 * from bhip_set_load, Newton with the tangent reassembled every iteration
 * until max|residual| < 2*max_allowable_residual or max_iterations extra
 * solves, then update("U").  opts->use_x0 = 0 starts every Newton correction
 * from zero; 1 starts it from the previous correction (the reference passes
 * the increment's displacement as the guess, equation.cpp:1678-1686). */
int bhip_plastic_increment(bhip_ctx *ctx, const double *presc_increment, double max_allowable_residual,
                           int max_iterations, const bhip_solve_opts *opts, bhip_plastic_stats *stats);
/* Converged state: stress (n_el*nip*6), effective plastic strain (n_el*nip),
 * accumulated displacement (n_dof); any pointer may be NULL. */
int bhip_plastic_get_state(bhip_ctx *ctx, double *stress, double *eps_bar, double *u_total);

/* ------------------------------------------------------ exchange formats */
#define BHIP_FMT_RCV_LOWER   1   /* "1", "L nnz", rows of "\t i j value" (lower, 0-based):
                                    SymmetricSparseMatrix::printMatrixAddressFormat (sparse.cpp:305-325) */
#define BHIP_FMT_RCV_UPPER   2   /* "2", upper rows: printMatrixAddressFormatUpper (sparse.cpp:266-290) */
#define BHIP_FMT_RCV         5   /* reading only: any RCV type 1-3 (SparseMatrixInterface::ReadMatrix,
                                    sparseMatrixInterface.cpp:51-101) */
#define BHIP_FMT_CSR_ASCII   3   /* writeCSRMatrixInASCII (SparseMatrixUtility.cpp:11-26) of the PARDISO
                                    storage (Pardiso.cpp:208-211): upper triangle, 1-based */
#define BHIP_FMT_CSR_BINARY  4   /* writeCSRMatrixInBinary (SparseMatrixUtility.cpp:28-38), same storage */

#define BHIP_TRI_LOWER       1
#define BHIP_TRI_UPPER       2
#define BHIP_TRI_FULL        3

#define BHIP_VEC_LOAD        1   /* LoadVector.txt (equation.cpp:1357-1362): one value per line */
#define BHIP_VEC_SOLUTION    2   /* Solution.txt (equation.cpp:1378-1383): count, then one value per line */

/* Write the context's matrix in a reference format (the printCSRFormat* /
 * printMatrixAddressFormat dumps of Equations::solve, equation.cpp:1345-1356). */
int bhip_write_matrix(bhip_ctx *ctx, const char *path, int format);
/* Write raw CSR arrays in a reference format (CSR formats: as given; RCV
 * formats: 0-based rows of the triangle to write).  No GPU needed. */
int bhip_write_csr_file(const char *path, int format, int n, long long nnz, const int *ia, const int *ja,
                        const double *a);
/* Load a matrix into the context instead of assembling one: a triangle
 * (expanded to the full symmetric matrix) or the full matrix, 0- or 1-based.
 * Every row needs its diagonal.  The context then solves with bhip_set_rhs +
 * bhip_solve / bhip_pcg_solve (FP32 inner values in CSR); mesh, assembly and
 * load calls report BHIP_ERR_STATE until bhip_set_mesh. */
int bhip_set_csr(bhip_ctx *ctx, int n, long long nnz, const int *ia, const int *ja, const double *a, int base,
                 int triangle);
/* Read a matrix file written by the reference (or by bhip_write_matrix) into
 * the context (readCSRMatrixInASCII/Binary, SparseMatrixUtility.cpp:40-78;
 * ReadMatrix for RCV); index base and triangle are taken from the file. */
int bhip_read_matrix(bhip_ctx *ctx, const char *path, int format);
/* LoadVector.txt / Solution.txt writers and readers (readLoadVector,
 * SparseMatrixUtility.cpp:79-85).  No GPU needed. */
int bhip_write_vector(const char *path, int kind, int n, const double *v);
int bhip_read_vector(const char *path, int kind, int n, double *v);

/* ------------------------------------------------------ transient diffusion */
#define BHIP_HEAT_X0_ZERO       0   /* initial guess du = 0 */
#define BHIP_HEAT_X0_PREVIOUS   1   /* du of the previous step (0 on the first) */
#define BHIP_HEAT_X0_REFERENCE  2   /* u_s on free DOFs, the prescribed value on restrained ones,
                                       as Equations::solve(u_s1) (equation.cpp:1367-1376) */
#define BHIP_HEAT_X0_PROJECTED  3   /* no reference counterpart: the Mbar-orthogonal projection of the
                                       step's solution on an Mbar-orthonormal basis of the increments
                                       since the last restart (up to 16),
                                       computed from the right-hand side alone (Fischer's projection
                                       for successive right-hand sides) */

typedef struct bhip_heat_stats {
    int    steps;            /* steps completed */
    int    iters;            /* PCG iterations over all steps (inner ones for FP32_IR) */
    int    outer_iters;      /* refinement steps over all steps */
    double max_rel_resid;    /* worst true relative residual of a step */
    double ms;               /* device time of the run (HIP events) */
} bhip_heat_stats;

/* DiffusionModel::DoTransientNonLinearAnalysis (DiffusionModel.cpp:711-831)
 * for a linear conductivity, after one BHIP_ASM_HEAT_TRANSIENT assembly with
 * coef = {alpha, dt} (Mbar = (a1 K + M)/dt, a1 = alpha dt): n_steps steps of
 * Mbar du = -K u_s + lift (assembleK_F_s_alpha_family, 322-347), u_s += du,
 * with the context's prescribed values as the increment of every restrained
 * DOF.  The reference reassembles K and M every step although they are
 * constant; here the right-hand side comes from the assembled K on the device.
 * u: n_dof, u_0 on entry, u after n_steps on return. */
int bhip_heat_run(bhip_ctx *ctx, int n_steps, double *u, int x0_mode, const bhip_solve_opts *opts,
                  bhip_heat_stats *stats);

/* --------------------------------------------------------------- serving */
/* No counterpart in the reference, which runs one model at a time on the
 * host.  The served workflow is the reference's macro-element / RVE one
 * (MultiFieldMacroElement::CalculateK, MultiFieldMacroElement.cpp:302-407):
 * assemble a hex8 elasticity mesh, then solve it for several prescribed
 * boundary-displacement load cases (bhip_solve_cases).  The server keeps up to
 * max_slots workers resident on one GPU, each with its own contexts and HIP
 * streams, starts queued requests in arrival order as soon as a worker is
 * free and the requests' estimated device bytes fit the budget (continuous
 * batching), packs small requests queued together into one packed solve
 * (bhip_pack_solve_cases), and reuses a context's pattern when its next
 * request has the same connectivity and restraints.  A request's result does
 * not depend on the other requests in flight or in its pack. */
typedef struct bhip_server bhip_server;

typedef struct bhip_server_opts {
    int    max_slots;    /* requests on the device at once (0: 4; at most 16) */
    double hbm_budget;   /* device bytes all slots may hold (0: 80% of free memory at creation) */
    double tol;          /* per case |b - K x| / |b| (0: 1e-8) */
    double inner_tol;    /* FP32_IR inner residual reduction (0: 1e-3) */
    int    maxit;        /* inner iteration cap per solve (0: 40000) */
    int    precision;    /* BHIP_PREC_FP32_IR (default), BHIP_PREC_FP64 or BHIP_PREC_FP64_BLOCKED */
    int    max_pack;     /* small requests solved together in one packed batched PCG
                            (0: 8; 1: no packing; at most 64) */
} bhip_server_opts;

/* One hex8 3D elasticity request (3 DOFs per node).  Every array is copied
 * by bhip_server_submit. */
typedef struct bhip_request {
    int n_nodes;
    const double *xyz;            /* 3 * n_nodes */
    int n_el;
    const int *conn;              /* 8 * n_el, reference node order */
    const unsigned char *rest;    /* 3 * n_nodes, nonzero = prescribed DOF */
    int mat_kind;                 /* BHIP_MAT_ISO_ELASTIC or BHIP_MAT_CMAT3D */
    int n_groups;
    const double *mat;            /* n_groups records */
    const int *elem_group;        /* n_el, or NULL (as bhip_set_materials) */
    int n_cases;                  /* load cases (prescribed-value sets) */
    const double *presc;          /* n_cases * 3 * n_nodes */
    const double *load;           /* 3 * n_nodes shared by all cases, or NULL */
    int order;                    /* Gauss points per direction (0: 2) */
} bhip_request;

typedef struct bhip_request_stats {
    int    status;           /* BHIP_OK or the failing call's status */
    int    slot;             /* slot that ran the request */
    int    reused_pattern;   /* 1 if the slot's pattern was reused */
    int    iters;            /* inner iterations over the case groups */
    int    outer_iters;
    double rel_resid;        /* worst true relative residual over the cases */
    double queue_ms;         /* submit -> start */
    double service_ms;       /* start -> done */
    double latency_ms;       /* submit -> done */
    double start_ms, end_ms; /* start and end, ms since the server was created */
    int    pack_size;        /* requests solved in the same packed solve (1: alone) */
    double setup_ms;         /* mesh, pattern, assembly and right-hand sides of this request */
    double solve_ms;         /* the (packed) solve it took part in */
    long long case_iters;    /* inner iterations summed over the request's load cases */
    char   message[256];     /* error text when status != BHIP_OK */
} bhip_request_stats;

int  bhip_server_create(int device, const bhip_server_opts *opts, bhip_server **out);
/* Stops the workers after the requests they are running; queued requests are
 * dropped.  No call may wait on the server while it is destroyed. */
void bhip_server_destroy(bhip_server *srv);
/* Queues a copy of the request; BHIP_ERR_CAPACITY if its estimated device
 * bytes exceed the budget. */
int  bhip_server_submit(bhip_server *srv, const bhip_request *req, long long *id);
int  bhip_server_poll(bhip_server *srv, long long id, int *done);
/* Blocks until request `id` is done, copies its solutions (n_cases * 3 *
 * n_nodes, case-major; x may be NULL) and stats, and forgets the request.
 * Returns the request's status. */
int  bhip_server_wait(bhip_server *srv, long long id, double *x, bhip_request_stats *stats);
double bhip_server_budget(const bhip_server *srv);
/* The device-byte estimate admission uses for a request of this size. */
double bhip_request_footprint(int n_nodes, int n_el, int n_cases);

#ifdef __cplusplus
}
#endif
#endif
