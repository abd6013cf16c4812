// Compiled C++ caller of the C-ABI, in the order the reference's classes
// call their own code (Problem::init, Solver::preallocate_matrix,
// Assembler::assemble_new_load_step / assemble_nonlinear, the Newton /
// KSP solve and store_solution; src/Problem.cpp:105-330, src/Solver.cpp:365-441,
// src/Assembler.cpp:208-304, 554-666, src/SolverNonlinearKSP.cpp:55-159).
// Only include/pgfem.h and include/pgfem_host.h are used: this is what an
// AssemblerHIP / solver on the reference side links against.
//
// Workload: the unit square, n x n QUAD4, one centred circle r = 0.2,
// perfectly bonded (matrix E 3.2 GPa nu 0.35, inclusion E 70 GPa nu 0.22),
// left edge fixed, right edge u_x = 0.01, FP64 linear solve.
//
//   pgf_driver host <n> <out>   host set-up only: sizes and the CSR pattern
//                               (no device needed) into <out>
//   pgf_driver solve <n> <out>  plus pgf_create, one K+f assembly, one Newton
//                               step to 1e-8 and the solution into <out>
//
// <out> is a small binary file: int64 header (n_rows, nnz, n_all, dim), then
// rowptr [n_rows+1] int64, colidx [nnz] int32, and in solve mode the K values
// [nnz] FP64, P [n_rows] FP64 (before the solve, at u = 0 with the prescribed
// values applied), and the nodal solution [dim][n_all] FP64.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/pgfem.h"
#include "../../../include/pgfem_host.h"

namespace {

int fail_host(const char* what) {
    std::fprintf(stderr, "pgf_driver: %s failed: %s\n", what, pgfh_last_error());
    return 2;
}

int fail_dev(const char* what, pgf_ctx* ctx) {
    std::fprintf(stderr, "pgf_driver: %s failed: %s\n", what, pgf_last_error(ctx));
    return 3;
}

template <typename T>
void put(std::FILE* f, const std::vector<T>& v) {
    if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 4 || (std::strcmp(argv[1], "host") && std::strcmp(argv[1], "solve"))) {
        std::fprintf(stderr, "usage: %s host|solve <n> <out>\n", argv[0]);
        return 1;
    }
    const bool solve = std::strcmp(argv[1], "solve") == 0;
    const int n = std::atoi(argv[2]);
    if (n < 2 || n > 4000) {
        std::fprintf(stderr, "pgf_driver: n must be in [2, 4000]\n");
        return 1;
    }

    // ---- Problem::init: mesh, inclusion, materials, boundary conditions ----
    const double incl[5] = {0.0, 0.5, 0.5, 0.0, 0.2};
    const int32_t incl_mat[1] = {1};
    const double E[2] = {3.2e9, 70e9}, nu[2] = {0.35, 0.22};
    const int32_t bc_set[3] = {PGFH_SET_LEFT, PGFH_SET_LEFT, PGFH_SET_RIGHT};
    const int32_t bc_node[3] = {0, 0, 0};
    const int32_t bc_dof[3] = {0, 1, 0};
    const double bc_val[3] = {0.0, 0.0, 0.01};
    pgfh_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.elem_kind = PGF_QUAD4;
    cfg.N[0] = n;
    cfg.N[1] = n;
    cfg.N[2] = 0;
    cfg.maxs[0] = 1.0;
    cfg.maxs[1] = 1.0;
    cfg.n_incl = 1;
    cfg.incl = incl;
    cfg.incl_mat = incl_mat;
    cfg.n_mat = 2;
    cfg.mat_E = E;
    cfg.mat_nu = nu;
    cfg.plane_strain = 1;
    cfg.tet_weight_scale = 1.0;
    cfg.n_bc = 3;
    cfg.bc_set = bc_set;
    cfg.bc_node = bc_node;
    cfg.bc_dof = bc_dof;
    cfg.bc_val = bc_val;
    pgfh_problem* prob = nullptr;
    if (pgfh_problem_create(&cfg, &prob)) return fail_host("pgfh_problem_create");
    pgfh_info info;
    if (pgfh_problem_info(prob, &info)) return fail_host("pgfh_problem_info");
    std::printf("problem: %lld nodes (%lld with enrichments), %lld elements (%lld cut), %lld free dofs, %lld prescribed\n",
                (long long)info.n_nodes, (long long)info.n_all_nodes, (long long)info.n_elem, (long long)info.n_cut,
                (long long)info.n_free, (long long)info.n_pres);

    // ---- Solver::preallocate_matrix: the CSR pattern ----
    int64_t n_rows = 0, nnz = 0;
    int32_t n_colors = 0;
    if (pgfh_pattern_info(prob, &n_rows, &nnz, &n_colors)) return fail_host("pgfh_pattern_info");
    std::vector<int64_t> rowptr(n_rows + 1);
    std::vector<int32_t> rowsplit(n_rows), colidx(nnz), color(info.n_elem);
    if (pgfh_get_pattern(prob, rowptr.data(), rowsplit.data(), colidx.data(), color.data()))
        return fail_host("pgfh_get_pattern");
    std::printf("pattern: %lld rows, %lld nonzeros, %d colours\n", (long long)n_rows, (long long)nnz, n_colors);

    std::FILE* out = std::fopen(argv[3], "wb");
    if (!out) {
        std::fprintf(stderr, "pgf_driver: cannot open %s\n", argv[3]);
        return 1;
    }
    const std::vector<int64_t> head = {n_rows, nnz, info.n_all_nodes, info.dim};
    put(out, head);
    put(out, rowptr);
    put(out, colidx);
    if (!solve) {
        std::fclose(out);
        pgfh_problem_destroy(prob);
        return 0;
    }

    // ---- device context: one upload of the mesh, tables and dofs ----
    pgf_mesh_desc desc;
    if (pgfh_mesh_desc(prob, &desc)) return fail_host("pgfh_mesh_desc");
    pgf_ctx* ctx = nullptr;
    if (pgf_create(&desc, &ctx)) {
        std::fprintf(stderr, "pgf_driver: pgf_create failed: %s\n", pgf_last_error(nullptr));
        return 3;
    }
    pgf_csr_info csr;
    if (pgf_csr_info_get(ctx, &csr)) return fail_dev("pgf_csr_info_get", ctx);
    if (csr.n_rows != n_rows || csr.nnz != nnz) {
        std::fprintf(stderr, "pgf_driver: device pattern differs from the host pattern\n");
        return 4;
    }
    std::vector<int64_t> drp(n_rows + 1);
    std::vector<int32_t> dsplit(n_rows), dcol(nnz);
    if (pgf_get_csr(ctx, drp.data(), dsplit.data(), dcol.data())) return fail_dev("pgf_get_csr", ctx);
    if (drp != rowptr || dcol != colidx || dsplit != rowsplit) {
        std::fprintf(stderr, "pgf_driver: device CSR differs from the host pattern\n");
        return 4;
    }

    // ---- Assembler::assemble_new_load_step: prescribed values, full load ----
    std::vector<int32_t> dof((size_t)info.dim * info.n_all_nodes);
    std::vector<double> pres(info.n_pres > 0 ? info.n_pres : 1);
    if (pgfh_get_dofs(prob, dof.data(), pres.data())) return fail_host("pgfh_get_dofs");
    if (pgf_set_load(ctx, pres.data(), nullptr)) return fail_dev("pgf_set_load", ctx);
    if (pgf_apply_load_fraction(ctx, 1.0)) return fail_dev("pgf_apply_load_fraction", ctx);

    // ---- Assembler::assemble_nonlinear: K and P on all rows ----
    if (pgf_assemble(ctx, 1.0, PGF_ASM_JAC | PGF_ASM_FUNC | PGF_ASM_ALL_ROWS)) return fail_dev("pgf_assemble", ctx);
    std::vector<double> K(nnz), P(n_rows);
    if (pgf_get_matrix(ctx, K.data())) return fail_dev("pgf_get_matrix", ctx);
    if (pgf_get_internal_force(ctx, P.data())) return fail_dev("pgf_get_internal_force", ctx);

    // ---- Newton step to ||r|| < 1e-8 ||r0|| (KSP: Jacobi-PCG) ----
    pgf_newton_opts o;
    std::memset(&o, 0, sizeof(o));
    o.rel_tol = 1e-8;
    o.max_iter = 40;
    o.precision = 64;
    o.lin_rtol = 1e-10;
    o.lin_maxit = 50000;
    o.dt = 1.0;
    pgf_step_info si;
    const int rc = pgf_newton_step(ctx, 1.0, &o, &si);
    if (rc) return fail_dev("pgf_newton_step", ctx);
    std::printf("newton: %d iterations, %d PCG iterations, |r| %.3e -> %.3e, converged %d\n", si.newton_iterations,
                si.pcg_iterations, si.res0, si.res, si.converged);
    if (!si.converged) return 5;

    // ---- store_solution ----
    std::vector<double> U((size_t)info.dim * info.n_all_nodes);
    if (pgf_get_solution(ctx, U.data())) return fail_dev("pgf_get_solution", ctx);
    put(out, K);
    put(out, P);
    put(out, U);
    std::fclose(out);
    pgf_destroy(ctx);
    pgfh_problem_destroy(prob);
    return 0;
}
