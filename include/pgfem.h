/*
 * pgfem.h — C-ABI of the MI355X assemble-and-solve path of PIGFEM.
 *
 * One context (pgf_ctx) holds one mesh resident in HBM: nodal coordinates,
 * element connectivity, cut-element (IGFEM) integration tables, the cohesive
 * internal state, the CSR stiffness matrix with the reference's sparsity
 * pattern, and the Krylov work vectors.  Every entry point returns 0 on success
 * or a negative status (PGF_E*); nothing exits the process.  Host buffers are
 * caller-owned and copied; device buffers are owned by the context.  One host
 * thread per context.  All device work runs on the context's own HIP stream.
 *
 * Reference interfaces replaced (paths relative to PIGFEM4.6.1/):
 *   pgf_create / pgf_get_csr       Solver::preallocate_matrix + populate_local_neighbors
 *                                  (src/Solver.cpp:365-441, 723-829) and
 *                                  DofObject::get_elem_global_dofs (src/DofObject.cpp:189-211)
 *   pgf_assemble                   Assembler::assemble_linear / assemble_nonlinear with their
 *                                  element loops assemble_elem_linear / assemble_elem_nonlinear
 *                                  (src/Assembler.cpp:48-93, 94-207, 208-304, 305-553,
 *                                  include/Assembler.h:41-51) and the scatter set_matrix / set_vec
 *                                  (src/Assembler.cpp:811-879)
 *                                  with KernelLinear / KernelNonlinear / KernelCohesive
 *                                  (src/AssemblerLinearElasticity.cpp:25-38,
 *                                  src/AssemblerSmallStrainNonlinear.cpp:28-91)
 *   pgf_set_solution               Solver::store_solution (src/Solver.cpp:860-973)
 *   pgf_set_load                   Assembler::assemble_new_load_step (src/Assembler.cpp:554-666)
 *   pgf_isv_commit / _rollback     InternalVars::copy_data_in / copy_data (src/InternalVars.cpp:43-86)
 *   pgf_spmv                       MatMult(K_ff, ...) (src/SolverLinear.cpp:128)
 *   pgf_pcg                        KSPSolve with KSPCG + PCJACOBI (the commented-out choice at
 *                                  src/Solver.cpp:80-91; the active MUMPS Cholesky at :71-79 has no
 *                                  device counterpart)
 *   pgf_newton_step                SolverNonlinearKSP::nonlinearStep (src/SolverNonlinearKSP.cpp:55-159)
 *                                  standing in for SNESSolve (src/SolverNonlinearSNES.cpp:182-241)
 */
#ifndef PGFEM_H
#define PGFEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGF_API __attribute__((visibility("default")))

/* status codes */
#define PGF_OK 0
#define PGF_EARG (-1)
#define PGF_EUNSUPPORTED (-2)
#define PGF_EGEOMETRY (-3)
#define PGF_EDEVICE (-4)
#define PGF_ENOTCONVERGED (-5)
#define PGF_ESTATE (-6)
#define PGF_ELIMIT (-7)

/* element kinds */
#define PGF_QUAD4 0
#define PGF_TET4 1

/* cohesive laws */
#define PGF_LAW_OP_EXPONENTIAL 0
#define PGF_LAW_BILINEAR 1

/* pgf_assemble flags (assembleFunc / assembleJac of the reference) */
#define PGF_ASM_FUNC 1u      /* internal force P_int, FP64 */
#define PGF_ASM_JAC 2u       /* stiffness K */
#define PGF_ASM_FP32 4u      /* K in FP32 (mixed-precision path), each entry the FP64 sum rounded once;
                                the FP64 matrix is scratch for it (read it after an FP64 assembly);
                                P_int stays FP64 */
#define PGF_ASM_ALL_ROWS 8u  /* also assemble prescribed rows (K_pf, K_pp); default free rows only */
#define PGF_ASM_COLORED 16u  /* element-coloured read-modify-write scheme instead of the row gather */

typedef struct pgf_ctx pgf_ctx;

/* Mesh as built on the host (Problem::init of the reference).  All arrays are
 * read during pgf_create only. */
typedef struct pgf_mesh_desc {
    int32_t dim;               /* 2 or 3 */
    int32_t elem_kind;         /* PGF_QUAD4 or PGF_TET4 */
    int32_t plane_strain;      /* 2D only */
    int32_t reserved0;
    int64_t n_nodes;           /* regular nodes */
    int64_t n_all_nodes;       /* regular + enrichment (+ mirror) nodes */
    const double* coords;      /* [3][n_all_nodes], structure of arrays */
    int64_t n_elem;
    const int32_t* conn;       /* [n_elem][4] regular nodes */
    const int32_t* elem_mat;   /* [n_elem] volumetric material of uncut elements */
    int32_t n_mat;
    int32_t n_coh_law;
    const double* mat_E;       /* [n_mat] */
    const double* mat_nu;      /* [n_mat] */
    const int32_t* coh_type;   /* [n_coh_law] PGF_LAW_* */
    const double* coh_params;  /* [n_coh_law][6]: sigma_c, delta_c, delta_f, beta, viscosity, t_floor */
    /* cut (intersected) elements */
    int64_t n_cut;
    const int32_t* cut_elem;      /* [n_cut] parent element */
    const int32_t* cut_n_enrich;  /* [n_cut] enrichment slots (incl. mirrors) */
    const int32_t* cut_enrich;    /* [n_cut][16] enrichment-node index per slot */
    const int64_t* vpt_ptr;       /* [n_cut+1] volumetric integration points per cut element */
    const double* vpt_w;          /* J*W */
    const int32_t* vpt_mat;
    const int64_t* vpt_grad_off;  /* offset of each point's gradient block */
    const double* vpt_grad;       /* per point [n_total][dim] */
    const int64_t* cpt_ptr;       /* [n_cut+1] cohesive integration points per cut element */
    const double* cpt_w;          /* dA*W */
    const double* cpt_N;          /* [n_cpt][4] */
    const double* cpt_R;          /* [n_cpt][9] rotation, rows: tangents then unit normal */
    const uint8_t* cpt_nodes;     /* [n_cpt][8] element slots: lower surface then upper surface */
    const uint8_t* cpt_nh;        /* [n_cpt] nodes per surface */
    const int32_t* cpt_law;       /* [n_cpt] */
    /* degrees of freedom */
    int64_t n_free;
    int64_t n_pres;
    const int32_t* dof;           /* [dim][n_all_nodes] global dof id (free first) */
} pgf_mesh_desc;

typedef struct pgf_csr_info {
    int64_t n_rows;   /* n_free + n_pres */
    int64_t n_free;
    int64_t nnz;      /* all four blocks */
    int64_t nnz_ff;
    int32_t n_colors;
    int32_t reserved;
} pgf_csr_info;

typedef struct pgf_solve_info {
    int32_t iterations;
    int32_t converged;
    double rel_residual;   /* ||b - A x|| / ||b|| as tracked by the recurrence */
    double ms;             /* wall time of the solve on the device stream */
} pgf_solve_info;

typedef struct pgf_newton_opts {
    double rel_tol;        /* ||r|| < rel_tol*||r0||  (reference rel_tol 1e-8) */
    double abs_tol;        /* or ||r|| < abs_tol */
    int32_t max_iter;      /* Newton iterations (reference max_iter 40) */
    int32_t precision;     /* 64: FP64 K and PCG; 32: FP32 K values, FP64 PCG vectors and residual */
    double lin_rtol;       /* PCG relative tolerance per Newton iteration */
    int32_t lin_maxit;
    int32_t use_guess;     /* 1: start PCG from the device guess buffer (pgf_set_guess); 2, 3: learned guess
                              (pgfem_learn.h) */
    double dt;             /* time increment passed to the cohesive law */
    int32_t adaptive_lin;  /* 1: per-iteration PCG tolerance max(lin_rtol, 0.5 rel_tol ||r0|| / ||r_k||)
                              (inexact Newton forcing term); 0: lin_rtol every iteration */
    int32_t fp32_maxit;    /* precision 32: PCG iterations allowed on the FP32 matrix before that solve
                              continues from its iterate on the FP64 matrix, and the rest of the step
                              runs in FP64 (0: lin_maxit) */
} pgf_newton_opts;

typedef struct pgf_step_info {
    int32_t newton_iterations;
    int32_t converged;
    int32_t pcg_iterations;     /* summed over the step */
    int32_t fp64_fallback;      /* precision 32: Newton iteration (1-based) whose FP32 solve did not
                                   converge and was finished in FP64, 0 if none */
    double res0;                /* ||r0||, FP64 */
    double res;                 /* final ||r||, FP64 */
    double ms_total;            /* device time of the step */
    double ms_assemble;
    double ms_solve;
} pgf_step_info;

/* ---- context ---- */
PGF_API int pgf_create(const pgf_mesh_desc* desc, pgf_ctx** out);
PGF_API void pgf_destroy(pgf_ctx* ctx);
/* last error message of ctx, or of the last failed pgf_create when ctx is NULL */
PGF_API const char* pgf_last_error(const pgf_ctx* ctx);
PGF_API int pgf_device_info(char* name, int32_t name_len, int32_t* n_cu, double* hbm_gb);

/* ---- pattern (integer, bit-exact with the reference) ---- */
PGF_API int pgf_csr_info_get(pgf_ctx* ctx, pgf_csr_info* info);
PGF_API int pgf_get_csr(pgf_ctx* ctx, int64_t* rowptr, int32_t* rowsplit, int32_t* colidx);
PGF_API int pgf_get_colors(pgf_ctx* ctx, int32_t* color);

/* ---- state ---- */
/* nodal displacement of every node, [dim][n_all_nodes] */
PGF_API int pgf_set_solution(pgf_ctx* ctx, const double* u_nodal);
PGF_API int pgf_get_solution(pgf_ctx* ctx, double* u_nodal);
/* nodal displacement from a dof vector: free part x_free [n_free], prescribed
 * part u_pres [n_pres] (NULL: zero); nodes sharing a dof (periodic images) all
 * take its value (the reference's store_solution, src/Solver.cpp:860-973) */
PGF_API int pgf_set_solution_dofs(pgf_ctx* ctx, const double* x_free, const double* u_pres);
/* prescribed values (full load) [n_pres] and external free-dof force [n_free] (may be NULL) */
PGF_API int pgf_set_load(pgf_ctx* ctx, const double* u_pres_full, const double* f_ext_full);
/* apply load fraction t/T: prescribed dofs of the nodal field <- frac*u_pres_full */
PGF_API int pgf_apply_load_fraction(pgf_ctx* ctx, double frac);

/* ---- assembly ---- */
PGF_API int pgf_assemble(pgf_ctx* ctx, double dt, uint32_t flags);
PGF_API int pgf_get_matrix(pgf_ctx* ctx, double* vals);        /* [nnz] FP64 values */
PGF_API int pgf_get_matrix32(pgf_ctx* ctx, float* vals);       /* [nnz] FP32 values */
PGF_API int pgf_get_internal_force(pgf_ctx* ctx, double* p);   /* [n_rows] */

/* ---- cohesive internal state ---- */
PGF_API int pgf_n_cohesive_points(pgf_ctx* ctx, int64_t* n);
PGF_API int pgf_isv_commit(pgf_ctx* ctx);     /* trial -> converged */
PGF_API int pgf_isv_rollback(pgf_ctx* ctx);   /* converged -> trial */
/* delta_max (trial and converged) and damage flags per cohesive point:
 * bit0 loading branch, bits1-2 damage zone, bit3 softening started, bit4 fully separated */
PGF_API int pgf_get_isv(pgf_ctx* ctx, double* delta_max_trial, double* delta_max_conv, uint8_t* flags);
PGF_API int pgf_set_isv(pgf_ctx* ctx, const double* delta_max_conv);

/* ---- Krylov ---- */
PGF_API int pgf_spmv(pgf_ctx* ctx, const double* x, double* y);   /* y = K_ff x, host vectors [n_free] */
PGF_API int pgf_pcg(pgf_ctx* ctx, const double* b, double* x, const double* x0, double rtol, int32_t maxit,
                    int32_t precision, pgf_solve_info* info);
/* optional initial guess for the next Newton solves, [n_free] (learned predictor) */
PGF_API int pgf_set_guess(pgf_ctx* ctx, const double* dx);

/* ---- batched solves of independent systems (RVEs x load cases) ----
 * Replaces the per-RVE KSPSolve of the subscale loop (src/SubscaleModel.cpp:42-196)
 * and the multi-right-hand-side solve of SensitivityDirect::solveProblem
 * (MatMatSolve, src/SensitivityDirect.cpp:96-136).  K_ff of the context must be
 * block diagonal over the free-row ranges sys_ptr[0..n_sys] (a context built on
 * a concatenation of independent meshes; checked).  Every (system, rhs) pair
 * runs its own Jacobi-PCG from a zero start to ||r|| <= rtol ||b||, all pairs
 * in one captured-graph iteration sweep.  b, x: host [nrhs][n_free], nrhs 1..4;
 * iters, relres: [n_sys][nrhs]; tune[0..1] in: threads per row (4/8/16) and rows
 * per tile, 0 = autotune; out: the choice. */
PGF_API int pgf_batch_pcg(pgf_ctx* ctx, int32_t n_sys, const int64_t* sys_ptr, int32_t nrhs, const double* b,
                          double* x, double rtol, int32_t maxit, int32_t* iters, double* relres, double* ms,
                          int32_t* tune);
/* The same solve as one persistent cooperative launch: every iteration of
 * every pair runs inside a single kernel, grid-wide barriers between the
 * phases (p update, SpMM, alpha, vector update, beta and convergence), no
 * host round trip and no graph replay until all pairs have converged or hit
 * maxit.  Same arithmetic in the same order: results equal pgf_batch_pcg's
 * bit for bit.  Same arguments.  Holds at most 256 row tiles, one block per
 * CU (PGF_EUNSUPPORTED beyond): the size it has been run and checked at. */
PGF_API int pgf_batch_pcg_persistent(pgf_ctx* ctx, int32_t n_sys, const int64_t* sys_ptr, int32_t nrhs,
                                     const double* b, double* x, double rtol, int32_t maxit, int32_t* iters,
                                     double* relres, double* ms, int32_t* tune);

/* right-hand sides of prescribed-displacement load cases: b = -K_fp u_p
 * (src/SolverLinear.cpp:128-129 per load case); up host [nrhs][n_pres], b host [nrhs][n_free] */
PGF_API int pgf_rhs_prescribed(pgf_ctx* ctx, int32_t nrhs, const double* up, double* b);
/* reactions R = K_pf u_f + K_pp u_p (src/SolverLinear.cpp:135-137); needs an
 * assembly with PGF_ASM_ALL_ROWS; uf [nrhs][n_free], up [nrhs][n_pres], R [nrhs][n_pres] */
PGF_API int pgf_reactions(pgf_ctx* ctx, int32_t nrhs, const double* uf, const double* up, double* R);

/* homogenised response of RVEs at the current nodal solution u (the periodic
 * fluctuation) and macro strain eps_bar [3 or 6] (Voigt, engineering shear):
 * for every element segment k = [seg_elem[k], seg_elem[k+1]) the volume averages
 * of sigma = D (eps(u) + eps_bar) and of eps(u) over its volumetric integration
 * points, and the volume sum J*W.  out [n_seg][2*S+1].  Replaces
 * SubscaleModel::homogenize_stress_strain (src/SubscaleModel.cpp:42-196). */
PGF_API int pgf_volume_average(pgf_ctx* ctx, const double* eps_bar, int32_t n_seg, const int64_t* seg_elem,
                               double* out);

/* Allocation-free variant of pgf_volume_average for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * No counterpart in the reference.  The staging of eps_bar and the segment
 * list, the per-segment results and the partial sums live in one static
 * device array per GPU (about 6.9 MB, held until the process ends), so
 * n_seg is limited to PGF_NOALLOC_MAX_SEG (PGF_ELIMIT beyond it).  Calls that
 * overlap on one GPU share that array and race; issue them from one host
 * thread.  Same results as pgf_volume_average, bit for bit. */
#define PGF_NOALLOC_MAX_SEG 1024
PGF_API int pgf_volume_average_noalloc(pgf_ctx* ctx, const double* eps_bar, int32_t n_seg, const int64_t* seg_elem,
                                       double* out);

/* ---- shape sensitivities (direct method, SensitivityDirect::solveProblem,
 * src/SensitivityDirect.cpp:25-188, with SensitivityShapeParameterStructural,
 * src/SensitivityShapeParameterStructural.cpp:18-61) ----
 * Tables from pgfh_shape_sensitivity (or their concatenation for a batch);
 * param [n_cut] is the parameter slot of each cut element (-1: none). */
PGF_API int pgf_set_shape_sensitivity(pgf_ctx* ctx, const double* dgrad, const double* divv, const int32_t* param);
/* At the current nodal solution and macro strain eps_bar [S] (zero for a
 * problem without body load): rhs [n_free] = -(dP_int/dr - dF_body/dr) of
 * parameter slot `slot` (the direct method's right-hand side, K du/dr = rhs);
 * dsig [n_seg][S + 1] = integral over each element segment of the explicit
 * stress derivative D dB u + sigma div v, then the derivative of the
 * segment's volume, the sum of J*W div v (the discrete volume moves with r
 * where a quadrature does not integrate the area exactly).  The
 * homogenised-stress sensitivity is (integral of D B du/dr + dsig[:S]) /
 * volume - sigma_bar dsig[S] / volume.  rhs may be NULL; n_seg may be 0. */
PGF_API int pgf_shape_pseudo_load(pgf_ctx* ctx, const double* eps_bar, int32_t slot, double* rhs, int32_t n_seg,
                                  const int64_t* seg_elem, double* dsig);

/* ---- Newton load step on the device (residual FP64) ---- */
PGF_API int pgf_newton_step(pgf_ctx* ctx, double frac, const pgf_newton_opts* opts, pgf_step_info* info);

/* ---- measurement ---- */
/* time `reps` calls of pgf_assemble(flags) with HIP events on the context stream */
PGF_API int pgf_time_assemble(pgf_ctx* ctx, uint32_t flags, int32_t reps, double* ms_per_call);
/* per-pass device time of the row-gather assembly (out[0] call, out[1] cut-element pass, out[2] gather pass) */
PGF_API int pgf_time_assemble_split(pgf_ctx* ctx, uint32_t flags, int32_t reps, double* out);
/* time `reps` K_ff SpMVs on device-resident FP64 vectors, FP64 or FP32 values */
PGF_API int pgf_time_spmv(pgf_ctx* ctx, int32_t precision, int32_t reps, double* ms_per_call);

/* ---- tuning of the row-gather assembly (no counterpart in the reference:
 * the reference assembles with MatSetValue, src/Assembler.cpp:811-832) ----
 * slots: threads that share one unit's incident elements, 1, 2, 4 or 8;
 * node_rows: 0 = a unit is one matrix row, 1 = a unit is all rows of a node.
 * Default: 1 slot, one row per unit in 2D; 2 slots in 3D.  Rebuilds the
 * node-block partition; the assembled values do not change (bit for bit). */
PGF_API int pgf_set_gather_layout(pgf_ctx* ctx, int32_t slots, int32_t node_rows);
PGF_API int pgf_get_gather_layout(pgf_ctx* ctx, int32_t* slots, int32_t* node_rows);

#ifdef __cplusplus
}
#endif
#endif
