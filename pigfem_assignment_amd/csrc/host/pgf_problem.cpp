// Host problem set-up and the load-step driver (pgfh_* C-ABI).
//
// Problem::init order (src/Problem.cpp:383-594): nodes detected, elements
// detected, enrichments added, shape functions stored, then dofs distributed
// (src/Problem.cpp:597-616 via BoundaryObject::apply_BCs and
// DofObject::distribute_dofs).  The load-step loop is SolverNonlinear::solve
// (src/SolverNonlinear.cpp:111-397) with its step controller update_time_step
// (:455-491): accept -> t += dt, grow dt x1.5 after a first-try convergence;
// reject -> restore the converged field and shrink dt x0.25; dt clamped to
// [min_dt, max_dt] and trimmed to the remaining time.  The per-step
// penetration pre-pass (:178-192) leaves delta_max unchanged and is skipped.
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/pgfem_host.h"
#include "pgf_dofs.hpp"
#include "pgf_mesh.hpp"
#include "pgf_sens.hpp"

namespace {
std::string g_host_error;
}

namespace pgf {
void set_host_error(const std::string& msg) { g_host_error = msg; }
void write_vtk(const Mesh& m, const std::string& path, const double* u, const double* eps_bar, double t);
}

struct pgfh_problem {
    pgf::Mesh mesh;
    pgf::DofMap dofs;
    std::vector<double> mat_E, mat_nu;
    std::vector<int32_t> coh_type;
    std::vector<double> coh_params;
    std::vector<int32_t> cut_elem, cut_ne, cut_en;
    std::vector<double> coords;  // [3][n_all]
    int plane_strain = 1;
    double setup_ms = 0.0;
    double phase_ms[4] = {0, 0, 0, 0};
    int device_preprocess = 0;
    int64_t n_enrich_initial = 0;
    pgf::Pattern pattern;
    bool have_pattern = false;
    void ensure_pattern() {
        if (have_pattern) return;
        pgf::MeshView mv = pgf::MeshView::of(mesh, cut_elem, cut_ne, cut_en);
        pattern.build(mv, dofs.dof.data(), dofs.n_free, dofs.n_pres);
        have_pattern = true;
    }
};

extern "C" {

PGF_API const char* pgfh_last_error(void) { return g_host_error.c_str(); }

PGF_API int pgfh_problem_create(const pgfh_config* cfg, pgfh_problem** out) {
    auto t0 = std::chrono::steady_clock::now();
    pgfh_problem* p = nullptr;
    try {
        if (!cfg || !out) throw pgf::Error(PGF_EARG, "null config");
        if (cfg->n_mat <= 0) throw pgf::Error(PGF_EARG, "at least one material is required");
        p = new pgfh_problem();
        pgf::MeshOptions opt;
        opt.cohesive = cfg->cohesive != 0;
        opt.plane_strain = cfg->plane_strain != 0;
        opt.sphere_zlim_as_reference = cfg->sphere_zlim_as_reference != 0;
        opt.tet_weight_scale = cfg->tet_weight_scale > 0 ? cfg->tet_weight_scale : 1.0;
        opt.default_coh_law = cfg->default_coh_law;
        if (opt.cohesive && cfg->n_coh_law <= 0) throw pgf::Error(PGF_EARG, "cohesive mesh without a cohesive law");
        p->mesh.set_options(opt);
        p->mesh.generate_mesh(cfg->elem_kind == PGF_TET4 ? pgf::TET4 : pgf::QUAD4, cfg->mins, cfg->maxs, cfg->N);
        if (cfg->periodic) p->mesh.build_periodic_nodesets();
        for (int i = 0; i < cfg->n_incl; ++i) {
            pgf::Inclusion inc;
            inc.type = (int)cfg->incl[i * 5 + 0];
            inc.c[0] = cfg->incl[i * 5 + 1];
            inc.c[1] = cfg->incl[i * 5 + 2];
            inc.c[2] = cfg->incl[i * 5 + 3];
            inc.r = cfg->incl[i * 5 + 4];
            inc.mat = cfg->incl_mat ? cfg->incl_mat[i] : 1;
            inc.coh_law = cfg->incl_coh_law ? cfg->incl_coh_law[i] : -1;
            if (inc.mat < 0 || inc.mat >= cfg->n_mat) throw pgf::Error(PGF_EARG, "inclusion material out of range");
            if (inc.r <= 0) throw pgf::Error(PGF_EARG, "inclusion radius must be positive");
            p->mesh.add_inclusion(inc);
        }
        if (cfg->device_preprocess) {
            pgf::device_preprocess(p->mesh, p->phase_ms);
            p->device_preprocess = 1;
        } else {
            auto ms = [] { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
            double a = ms();
            p->mesh.analyze_nodes();
            double b = ms();
            p->mesh.analyze_elements();
            double c = ms();
            p->mesh.add_enrichments();
            double d = ms();
            p->mesh.store_shape_functions();
            double e = ms();
            p->phase_ms[0] = b - a; p->phase_ms[1] = c - b; p->phase_ms[2] = d - c; p->phase_ms[3] = e - d;
        }
        p->n_enrich_initial = opt.cohesive ? p->mesh.n_enrich_nodes() / 2 : p->mesh.n_enrich_nodes();
        // Dirichlet conditions (BoundaryObject::set_dirichlet_bcs_from_nodeset)
        std::vector<pgf::Dirichlet> bcs;
        static const char* names[] = {"LEFT", "RIGHT", "BOTTOM", "TOP", "FRONT", "BACK"};
        for (int i = 0; i < cfg->n_bc; ++i) {
            int set = cfg->bc_set[i];
            if (set == PGFH_SET_NODE) {
                bcs.push_back({cfg->bc_node[i], cfg->bc_dof[i], cfg->bc_val[i]});
            } else {
                if (set < 0 || set > 5) throw pgf::Error(PGF_EARG, "unknown nodeset code");
                for (int32_t n : p->mesh.nodeset(names[set])) bcs.push_back({n, cfg->bc_dof[i], cfg->bc_val[i]});
            }
        }
        p->dofs.build(p->mesh, bcs);
        p->mat_E.assign(cfg->mat_E, cfg->mat_E + cfg->n_mat);
        p->mat_nu.assign(cfg->mat_nu, cfg->mat_nu + cfg->n_mat);
        if (cfg->n_coh_law > 0) {
            p->coh_type.assign(cfg->coh_type, cfg->coh_type + cfg->n_coh_law);
            p->coh_params.assign(cfg->coh_params, cfg->coh_params + 6 * cfg->n_coh_law);
        }
        p->plane_strain = cfg->plane_strain;
        pgf::MeshView::of(p->mesh, p->cut_elem, p->cut_ne, p->cut_en);
        const int64_t na = p->mesh.n_all_nodes();
        p->coords.resize(3 * na);
        for (int d = 0; d < 3; ++d) std::memcpy(&p->coords[d * na], p->mesh.X(d).data(), sizeof(double) * na);
        p->setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        *out = p;
        return 0;
    } catch (const pgf::Error& e) {
        g_host_error = e.what();
        delete p;
        return e.code;
    } catch (const std::exception& e) {
        g_host_error = e.what();
        delete p;
        return PGF_EARG;
    }
}

PGF_API void pgfh_problem_destroy(pgfh_problem* p) { delete p; }

PGF_API int pgfh_problem_info(pgfh_problem* p, pgfh_info* info) {
    if (!p || !info) return PGF_EARG;
    const auto& m = p->mesh;
    info->dim = m.dim();
    info->n_colors_hint = 0;
    info->n_nodes = m.n_nodes();
    info->n_all_nodes = m.n_all_nodes();
    info->n_elem = m.n_elem();
    info->n_cut = (int64_t)m.cut().size();
    info->n_free = p->dofs.n_free;
    info->n_pres = p->dofs.n_pres;
    info->n_vpt = (int64_t)m.vpt_w.size();
    info->n_vgrad = (int64_t)m.vpt_grad.size();
    info->n_cpt = (int64_t)m.cpt_w.size();
    info->n_moved_nodes = m.n_moved_nodes();
    info->n_enrich_initial = p->n_enrich_initial;
    info->setup_ms = p->setup_ms;
    for (int k = 0; k < 4; ++k) info->phase_ms[k] = p->phase_ms[k];
    info->device_preprocess = p->device_preprocess;
    return 0;
}

PGF_API int pgfh_mesh_desc(pgfh_problem* p, pgf_mesh_desc* d) {
    if (!p || !d) return PGF_EARG;
    auto& m = p->mesh;
    std::memset(d, 0, sizeof(*d));
    d->dim = m.dim();
    d->elem_kind = m.kind() == pgf::QUAD4 ? PGF_QUAD4 : PGF_TET4;
    d->plane_strain = p->plane_strain;
    d->n_nodes = m.n_nodes();
    d->n_all_nodes = m.n_all_nodes();
    d->coords = p->coords.data();
    d->n_elem = m.n_elem();
    d->conn = m.conn().data();
    d->elem_mat = m.elem_mat().data();
    d->n_mat = (int32_t)p->mat_E.size();
    d->mat_E = p->mat_E.data();
    d->mat_nu = p->mat_nu.data();
    d->n_coh_law = (int32_t)p->coh_type.size();
    d->coh_type = p->coh_type.data();
    d->coh_params = p->coh_params.data();
    d->n_cut = (int64_t)m.cut().size();
    d->cut_elem = p->cut_elem.data();
    d->cut_n_enrich = p->cut_ne.data();
    d->cut_enrich = p->cut_en.data();
    d->vpt_ptr = m.vpt_ptr.data();
    d->vpt_w = m.vpt_w.data();
    d->vpt_mat = m.vpt_mat.data();
    d->vpt_grad_off = m.vpt_grad_off.data();
    d->vpt_grad = m.vpt_grad.data();
    d->cpt_ptr = m.cpt_ptr.data();
    d->cpt_w = m.cpt_w.data();
    d->cpt_N = m.cpt_N.data();
    d->cpt_R = m.cpt_R.data();
    d->cpt_nodes = m.cpt_nodes.data();
    d->cpt_nh = m.cpt_nh.data();
    d->cpt_law = m.cpt_law.data();
    d->n_free = p->dofs.n_free;
    d->n_pres = p->dofs.n_pres;
    d->dof = p->dofs.dof.data();
    return 0;
}

PGF_API int pgfh_get_coords(pgfh_problem* p, double* xyz) {
    if (!p || !xyz) return PGF_EARG;
    std::memcpy(xyz, p->coords.data(), sizeof(double) * p->coords.size());
    return 0;
}
PGF_API int pgfh_get_conn(pgfh_problem* p, int32_t* conn) {
    if (!p || !conn) return PGF_EARG;
    std::memcpy(conn, p->mesh.conn().data(), sizeof(int32_t) * p->mesh.conn().size());
    return 0;
}
PGF_API int pgfh_get_elem_mat(pgfh_problem* p, int32_t* mat) {
    if (!p || !mat) return PGF_EARG;
    std::memcpy(mat, p->mesh.elem_mat().data(), sizeof(int32_t) * p->mesh.elem_mat().size());
    return 0;
}
PGF_API int pgfh_get_node_detect(pgfh_problem* p, int32_t* det) {
    if (!p || !det) return PGF_EARG;
    std::memcpy(det, p->mesh.node_detect().data(), sizeof(int32_t) * p->mesh.node_detect().size());
    return 0;
}
PGF_API int pgfh_get_dofs(pgfh_problem* p, int32_t* dof, double* pres) {
    if (!p) return PGF_EARG;
    if (dof) std::memcpy(dof, p->dofs.dof.data(), sizeof(int32_t) * p->dofs.dof.size());
    if (pres) std::memcpy(pres, p->dofs.pres_value.data(), sizeof(double) * p->dofs.pres_value.size());
    return 0;
}
PGF_API int pgfh_get_enrich_edges(pgfh_problem* p, int32_t* e) {
    if (!p || !e) return PGF_EARG;
    const auto& ee = p->mesh.enrich_edge();
    for (size_t k = 0; k < ee.size(); ++k)
        for (int j = 0; j < 3; ++j) e[k * 3 + j] = ee[k][j];
    return 0;
}
PGF_API int pgfh_get_cut(pgfh_problem* p, int32_t* rec) {
    if (!p || !rec) return PGF_EARG;
    const auto& cut = p->mesh.cut();
    for (size_t c = 0; c < cut.size(); ++c) {
        int32_t* r = rec + c * 128;
        std::memset(r, 0, sizeof(int32_t) * 128);
        const auto& ce = cut[c];
        r[0] = ce.elem;
        r[1] = ce.n_cut;
        for (int i = 0; i < 8; ++i) { r[2 + i] = ce.cut_edge[i]; r[10 + i] = ce.enrich_incl[i]; }
        r[18] = ce.n_sub;
        for (int i = 0; i < 6; ++i) {
            r[19 + i] = ce.sub_nn[i];
            for (int k = 0; k < 8; ++k) r[25 + i * 8 + k] = ce.sub_nodes[i][k];
            r[73 + i] = ce.sub_mat[i];
        }
        r[79] = ce.n_coh;
        for (int i = 0; i < 2; ++i) {
            r[80 + i] = ce.coh_nn[i];
            for (int k = 0; k < 8; ++k) r[82 + i * 8 + k] = ce.coh_nodes[i][k];
            r[98 + i] = ce.coh_law[i];
        }
        r[100] = ce.n_enrich;
        for (int i = 0; i < 16; ++i) r[101 + i] = ce.enrich[i];
    }
    return 0;
}

PGF_API int pgfh_shape_sensitivity(pgfh_problem* p, int32_t as_reference, double* dgrad, double* divv, int32_t* param,
                                   double* vel) {
    if (!p || !dgrad || !divv || !param) return PGF_EARG;
    try {
        if (p->mesh.is_cohesive()) throw pgf::Error(PGF_EUNSUPPORTED, "shape sensitivities of cohesive meshes are not built");
        pgf::ShapeSensTables t;
        pgf::build_shape_sensitivity(p->mesh, as_reference != 0, t);
        std::memcpy(dgrad, t.dgrad.data(), sizeof(double) * t.dgrad.size());
        std::memcpy(divv, t.divv.data(), sizeof(double) * t.divv.size());
        std::memcpy(param, t.param.data(), sizeof(int32_t) * t.param.size());
        if (vel) std::memcpy(vel, t.vel.data(), sizeof(double) * t.vel.size());
        return 0;
    } catch (const pgf::Error& e) {
        g_host_error = e.what();
        return e.code;
    }
}

PGF_API int pgfh_write_vtk(pgfh_problem* p, const char* path, const double* u_nodal, const double* eps_bar, double t) {
    if (!p || !path) return PGF_EARG;
    try {
        pgf::write_vtk(p->mesh, path, u_nodal, eps_bar, t);
        return 0;
    } catch (const pgf::Error& e) {
        g_host_error = e.what();
        return e.code;
    }
}

PGF_API int pgfh_pattern_info(pgfh_problem* p, int64_t* n_rows, int64_t* nnz, int32_t* n_colors) {
    if (!p) return PGF_EARG;
    try {
        p->ensure_pattern();
    } catch (const pgf::Error& e) {
        g_host_error = e.what();
        return e.code;
    }
    if (n_rows) *n_rows = p->pattern.n_rows;
    if (nnz) *nnz = p->pattern.nnz;
    if (n_colors) *n_colors = p->pattern.n_colors;
    return 0;
}

PGF_API int pgfh_get_pattern(pgfh_problem* p, int64_t* rowptr, int32_t* rowsplit, int32_t* colidx, int32_t* color) {
    if (!p || !p->have_pattern) return PGF_ESTATE;
    const auto& P = p->pattern;
    if (rowptr) std::memcpy(rowptr, P.rowptr.data(), sizeof(int64_t) * P.rowptr.size());
    if (rowsplit) std::memcpy(rowsplit, P.rowsplit.data(), sizeof(int32_t) * P.rowsplit.size());
    if (colidx) std::memcpy(colidx, P.colidx.data(), sizeof(int32_t) * P.colidx.size());
    if (color) std::memcpy(color, P.color.data(), sizeof(int32_t) * P.color.size());
    return 0;
}

PGF_API int pgfh_solve_load_steps(pgfh_problem* p, pgf_ctx* ctx, const pgfh_load_opts* o, pgfh_step_cb cb, void* user,
                                  pgfh_load_info* info) {
    if (!p || !ctx || !o) return PGF_EARG;
    int rc = pgf_set_load(ctx, p->dofs.pres_value.data(), nullptr);
    if (rc) return rc;
    const int64_t nu = (int64_t)p->mesh.dim() * p->mesh.n_all_nodes();
    std::vector<double> Uprev(nu, 0.0);
    if ((rc = pgf_get_solution(ctx, Uprev.data()))) return rc;
    double t = 0.0, dt = o->max_time_step;
    const double T = o->T_final;
    pgfh_load_info li{};
    int steps = 0;
    while (t < T) {
        int repeat = 0;
        bool converged = false;
        while (!converged) {
            if (repeat >= o->max_repeats || dt < o->min_time_step) {
                g_host_error = "Newton did not converge within the load-step repeat limit";
                if (info) *info = li;
                return PGF_ENOTCONVERGED;
            }
            ++repeat;
            pgf_newton_opts no = o->newton;
            no.dt = dt;
            pgf_step_info si{};
            rc = pgf_newton_step(ctx, (t + dt) / T, &no, &si);
            if (rc != 0 && rc != PGF_ENOTCONVERGED) return rc;
            converged = rc == 0;
            li.newton_iterations += si.newton_iterations;
            li.pcg_iterations += si.pcg_iterations;
            li.ms_total += si.ms_total;
            li.ms_assemble += si.ms_assemble;
            li.ms_solve += si.ms_solve;
            if (converged) {
                t += dt;
                if (cb) cb(user, steps, t, dt, &si);
            } else {
                ++li.repeats;
                if ((rc = pgf_set_solution(ctx, Uprev.data()))) return rc;
                if ((rc = pgf_isv_rollback(ctx))) return rc;
            }
            // update_time_step
            if (!converged) dt *= 0.25;
            else if (repeat == 1) dt *= 1.5;
            double remaining = T - t;
            if (dt > o->max_time_step) dt = o->max_time_step;
            if (dt > remaining) {
                if (remaining >= o->min_time_step) dt = remaining;
                else t = T;
            }
        }
        if ((rc = pgf_isv_commit(ctx))) return rc;
        if ((rc = pgf_get_solution(ctx, Uprev.data()))) return rc;
        ++steps;
        if (o->max_steps > 0 && steps >= o->max_steps) break;
    }
    li.steps = steps;
    li.t_reached = t;
    if (info) *info = li;
    return 0;
}

}  // extern "C"
