/*
 * pgfem_host.h — host-side problem construction for the MI355X path.
 *
 * Mirrors the reference's set-up surface: Mesh::generate_mesh, add_inclusion,
 * material assignment, BoundaryObject::set_dirichlet_bcs_from_nodeset and
 * Problem::init (src/Problem.cpp:383-594: analyze_nodes, analyze_elements,
 * add_enrichments, store_shape_functions, DofObject::distribute_dofs).  The
 * result is exported as plain arrays and as a pgf_mesh_desc for pgf_create.
 * The load-step driver mirrors SolverNonlinear::solve / update_time_step
 * (src/SolverNonlinear.cpp:111-397, 455-491) on top of pgf_newton_step.
 */
#ifndef PGFEM_HOST_H
#define PGFEM_HOST_H

#include <stdint.h>

#include "pgfem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pgfh_problem pgfh_problem;

/* nodeset codes for Dirichlet conditions */
#define PGFH_SET_NODE (-1) /* bc_node holds a node id */
#define PGFH_SET_LEFT 0
#define PGFH_SET_RIGHT 1
#define PGFH_SET_BOTTOM 2
#define PGFH_SET_TOP 3
#define PGFH_SET_FRONT 4
#define PGFH_SET_BACK 5

typedef struct pgfh_config {
    int32_t elem_kind;          /* PGF_QUAD4 or PGF_TET4 */
    int32_t N[3];               /* divisions */
    double mins[3], maxs[3];
    int32_t n_incl;
    const double* incl;         /* [n_incl][5]: type (0 circle, 1 sphere), cx, cy, cz, r */
    const int32_t* incl_mat;    /* [n_incl] volumetric material */
    const int32_t* incl_coh_law;/* [n_incl] or NULL: cohesive law per inclusion (-1 = default) */
    int32_t n_mat;
    const double* mat_E;        /* material 0 is the matrix */
    const double* mat_nu;
    int32_t plane_strain;
    int32_t cohesive;
    int32_t n_coh_law;
    const int32_t* coh_type;
    const double* coh_params;   /* [n_coh_law][6] */
    int32_t default_coh_law;
    int32_t sphere_zlim_as_reference;
    double tet_weight_scale;    /* 1.0 reference, 1/6 true volumes; 0 means 1.0 */
    int32_t n_bc;
    const int32_t* bc_set;      /* PGFH_SET_* */
    const int32_t* bc_node;     /* node id when bc_set == PGFH_SET_NODE */
    const int32_t* bc_dof;
    const double* bc_val;
    /* periodic RVE (Mesh::build_periodic_nodesets, src/Mesh.cpp:603-741): the
     * nodes on the upper faces share the dofs of their images on the lower
     * faces, corners share node 0's; only the primary node's Dirichlet
     * conditions count (DofObject::set_local_owned_normal_dofs,
     * src/DofObject.cpp:336-386) */
    int32_t periodic;
    /* 1: node detection, element detection, enrichment nodes and the cut
     * elements' integration tables run on the current HIP device (same
     * results as the host, bit for bit); 0: on the host */
    int32_t device_preprocess;
} pgfh_config;

typedef struct pgfh_info {
    int32_t dim;
    int32_t n_colors_hint;
    int64_t n_nodes, n_all_nodes, n_elem, n_cut, n_free, n_pres;
    int64_t n_vpt, n_vgrad, n_cpt, n_moved_nodes, n_enrich_initial;
    double setup_ms;
    /* wall time of the IGFEM phases: node detection, element detection,
     * enrichment nodes, integration tables (device transfers included) */
    double phase_ms[4];
    int32_t device_preprocess;  /* 1 when the phases ran on the device */
} pgfh_info;

PGF_API int pgfh_problem_create(const pgfh_config* cfg, pgfh_problem** out);
PGF_API void pgfh_problem_destroy(pgfh_problem* p);
PGF_API const char* pgfh_last_error(void);
PGF_API int pgfh_problem_info(pgfh_problem* p, pgfh_info* info);
/* descriptor pointing into the problem's storage (valid while p lives) */
PGF_API int pgfh_mesh_desc(pgfh_problem* p, pgf_mesh_desc* d);

/* exports (caller-allocated, sizes from pgfh_info) */
PGF_API int pgfh_get_coords(pgfh_problem* p, double* xyz);              /* [3][n_all_nodes] */
PGF_API int pgfh_get_conn(pgfh_problem* p, int32_t* conn);              /* [n_elem][4] */
PGF_API int pgfh_get_elem_mat(pgfh_problem* p, int32_t* mat);           /* [n_elem] */
PGF_API int pgfh_get_node_detect(pgfh_problem* p, int32_t* det);        /* [n_nodes] */
PGF_API int pgfh_get_dofs(pgfh_problem* p, int32_t* dof, double* pres); /* [dim][n_all], [n_pres] */
/* enrichment nodes: [n_all-n_nodes][3] = (edge node 0, edge node 1, inclusion) */
PGF_API int pgfh_get_enrich_edges(pgfh_problem* p, int32_t* e);
/* cut-element topology, 128 int32 per cut element:
 *  [0] elem, [1] n_cut_edges, [2..9] cut edges, [10..17] inclusion per cut edge,
 *  [18] n_sub, [19..24] sub node counts, [25..72] sub nodes (6x8), [73..78] sub materials,
 *  [79] n_coh, [80..81] coh node counts, [82..97] coh nodes (2x8), [98..99] coh laws,
 *  [100] n_enrich, [101..116] enrichment node per slot */
PGF_API int pgfh_get_cut(pgfh_problem* p, int32_t* rec);

/* shape-sensitivity tables with respect to the radius of the inclusion that
 * cuts each element (SensitivityShapeParameter::compute_velocity and
 * Elem::ShapeFunctionPartials, src/SensitivityShapeParameter.cpp:49-182,
 * src/elem.cpp:141-292): dgrad [n_vgrad] derivative of every volumetric
 * point's slot gradients (layout of the descriptor's vpt_grad), divv [n_vpt]
 * velocity divergence, param [n_cut] inclusion of each cut element, vel
 * [n_all - n_nodes][3] (may be NULL) enrichment-node velocities.  as_reference
 * selects the reference's expression for the parent-slot term (see
 * pgf_sens.cpp) instead of the exact derivative.  Not built for cohesive
 * meshes. */
PGF_API int pgfh_shape_sensitivity(pgfh_problem* p, int32_t as_reference, double* dgrad, double* divv, int32_t* param,
                                   double* vel);

/* legacy ASCII VTK file of the mesh (cut elements as their integration
 * sub-elements), the displacement at every node (the physical value at
 * enrichment nodes, plus eps_bar x when eps_bar [3 or 6] is given) and the
 * material of every cell (Writer_ASCII_VTK / Writer_VTK,
 * src/Writer_ASCII_VTK.cpp:43-282, src/Writer_VTK.cpp:221-335, 639-730).
 * u_nodal [dim][n_all_nodes] may be NULL (mesh only). */
PGF_API int pgfh_write_vtk(pgfh_problem* p, const char* path, const double* u_nodal, const double* eps_bar, double t);

/* ---- batches of RVEs ----
 * Independent problems stacked into one block-diagonal problem for one
 * pgf_create and pgf_batch_pcg.  The reference has no batch: it solves each
 * RVE as its own Problem (SubscaleModel::assemble / solve,
 * src/SubscaleModel.cpp:42-196, once per macro point).  Regular nodes of all
 * parts come first (part order), then all enrichment nodes; part k's free
 * dofs are [free_ptr[k], free_ptr[k+1]) (its segment in pgf_batch_pcg), its
 * prescribed dofs n_free + [pres_ptr[k], pres_ptr[k+1]).  Parts must share
 * dimension, element kind and materials; cohesive parts are refused
 * (PGF_EUNSUPPORTED).  The parts may be destroyed after pgfh_batch_create. */
typedef struct pgfh_batch pgfh_batch;
PGF_API int pgfh_batch_create(pgfh_problem* const* parts, int32_t n_parts, pgfh_batch** out);
PGF_API void pgfh_batch_destroy(pgfh_batch* b);
/* descriptor pointing into the batch's storage (valid while b lives) */
PGF_API int pgfh_batch_desc(pgfh_batch* b, pgf_mesh_desc* d);
/* [n_parts+1] offsets each (any may be NULL): free dofs, prescribed dofs,
 * elements, regular nodes, enrichment nodes */
PGF_API int pgfh_batch_offsets(pgfh_batch* b, int64_t* free_ptr, int64_t* pres_ptr, int64_t* elem_ptr,
                               int64_t* reg_ptr, int64_t* enr_ptr);
/* [dim][n_all_nodes] batch dof ids, [n_pres] prescribed values */
PGF_API int pgfh_batch_get_dofs(pgfh_batch* b, int32_t* dof, double* pres);

/* CSR pattern and element colouring exactly as pgf_create builds them (host
 * integer work, no device needed): sizes first, then the arrays */
PGF_API int pgfh_pattern_info(pgfh_problem* p, int64_t* n_rows, int64_t* nnz, int32_t* n_colors);
PGF_API int pgfh_get_pattern(pgfh_problem* p, int64_t* rowptr, int32_t* rowsplit, int32_t* colidx, int32_t* color);

/* ---- drivers ---- */
typedef struct pgfh_load_opts {
    double T_final;          /* reference T_final */
    double max_time_step;
    double min_time_step;
    int32_t max_repeats;     /* 20 in the reference */
    int32_t max_steps;       /* stop after this many accepted steps (<=0: run to T_final) */
    pgf_newton_opts newton;
} pgfh_load_opts;

typedef struct pgfh_load_info {
    int32_t steps;            /* accepted load steps */
    int32_t repeats;          /* rejected attempts */
    int32_t newton_iterations;
    int32_t pcg_iterations;
    double t_reached;
    double ms_total;
    double ms_assemble;
    double ms_solve;
} pgfh_load_info;

/* per-step record: step index, t, dt, newton its, pcg its, res0, res, ms */
typedef void (*pgfh_step_cb)(void* user, int32_t step, double t, double dt, const pgf_step_info* si);

PGF_API int pgfh_solve_load_steps(pgfh_problem* p, pgf_ctx* ctx, const pgfh_load_opts* o, pgfh_step_cb cb, void* user,
                                  pgfh_load_info* info);

#ifdef __cplusplus
}
#endif
#endif
