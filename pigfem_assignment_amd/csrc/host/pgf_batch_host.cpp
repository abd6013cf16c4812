// Batch of independent RVE problems as one block-diagonal problem (pgfh_batch_*).
//
// The reference solves its RVEs one at a time, each as its own Problem
// (SubscaleModel::assemble / solve, src/SubscaleModel.cpp:42-196, called per
// macro point).  Here the RVEs of a batch are stacked into one mesh
// descriptor so that one pgf_create holds them all and pgf_batch_pcg solves
// them together:
//   * regular nodes of every part first (part order), then every part's
//     enrichment nodes, so that [0, n_nodes) stays the regular-node range the
//     kernels assume;
//   * elements, cut elements and integration points in part order, with node,
//     element and table offsets added;
//   * free dofs of part k renumbered into [free_ptr[k], free_ptr[k+1]) and its
//     prescribed dofs into n_free + [pres_ptr[k], pres_ptr[k+1]), so each
//     RVE's free block is contiguous (the segments of pgf_batch_pcg).
// Parts must share dimension, element kind, plane-strain flag and materials;
// cohesive parts are refused (the batched solve is tested on linear RVEs).
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/pgfem_host.h"
#include "pgf_mesh.hpp"

struct pgfh_batch {
    int32_t dim = 0, elem_kind = 0, plane_strain = 0;
    int64_t n_nodes = 0, n_all = 0, n_elem = 0, n_cut = 0, n_free = 0, n_pres = 0;
    std::vector<double> coords, mat_E, mat_nu, vpt_w, vpt_grad, pres;
    std::vector<int32_t> conn, emat, cut_elem, cut_ne, cut_en, vpt_mat, dof;
    std::vector<int64_t> vpt_ptr, vpt_goff, cpt_ptr;
    std::vector<int64_t> free_ptr, pres_ptr, elem_ptr, reg_ptr, enr_ptr;
};

namespace pgf {
void set_host_error(const std::string& msg);  // read back by pgfh_last_error
}

namespace {
int fail(int code, const std::string& msg) {
    pgf::set_host_error(msg);
    return code;
}
}  // namespace

extern "C" {

PGF_API int pgfh_batch_create(pgfh_problem* const* parts, int32_t n_parts, pgfh_batch** out) {
    if (!parts || !out || n_parts <= 0) return fail(PGF_EARG, "pgfh_batch_create: no parts");
    *out = nullptr;
    std::vector<pgf_mesh_desc> d(n_parts);
    std::vector<std::vector<double>> pres(n_parts);
    for (int k = 0; k < n_parts; ++k) {
        if (!parts[k] || pgfh_mesh_desc(parts[k], &d[k]) != 0) return fail(PGF_EARG, "pgfh_batch_create: null part");
        pres[k].resize((size_t)d[k].n_pres);
        if (pgfh_get_dofs(parts[k], nullptr, pres[k].data()) != 0) return fail(PGF_EARG, "pgfh_batch_create: dofs");
        const pgf_mesh_desc& a = d[0];
        const pgf_mesh_desc& b = d[k];
        if (b.dim != a.dim || b.elem_kind != a.elem_kind || b.plane_strain != a.plane_strain || b.n_mat != a.n_mat)
            return fail(PGF_EARG, "pgfh_batch_create: part " + std::to_string(k) + " differs in element kind or materials");
        for (int i = 0; i < a.n_mat; ++i)
            if (b.mat_E[i] != a.mat_E[i] || b.mat_nu[i] != a.mat_nu[i])
                return fail(PGF_EARG, "pgfh_batch_create: part " + std::to_string(k) + " has other material constants");
        if (b.n_cut > 0 && b.cpt_ptr[b.n_cut] > 0)
            return fail(PGF_EUNSUPPORTED, "pgfh_batch_create: cohesive RVEs are not batched");
    }
    pgfh_batch* B = new pgfh_batch();
    const int dim = d[0].dim;
    B->dim = dim;
    B->elem_kind = d[0].elem_kind;
    B->plane_strain = d[0].plane_strain;
    B->mat_E.assign(d[0].mat_E, d[0].mat_E + d[0].n_mat);
    B->mat_nu.assign(d[0].mat_nu, d[0].mat_nu + d[0].n_mat);
    auto offsets = [&](auto get) {
        std::vector<int64_t> o(n_parts + 1, 0);
        for (int k = 0; k < n_parts; ++k) o[k + 1] = o[k] + get(d[k]);
        return o;
    };
    B->reg_ptr = offsets([](const pgf_mesh_desc& x) { return x.n_nodes; });
    B->enr_ptr = offsets([](const pgf_mesh_desc& x) { return x.n_all_nodes - x.n_nodes; });
    B->elem_ptr = offsets([](const pgf_mesh_desc& x) { return x.n_elem; });
    B->free_ptr = offsets([](const pgf_mesh_desc& x) { return x.n_free; });
    B->pres_ptr = offsets([](const pgf_mesh_desc& x) { return x.n_pres; });
    const int64_t NR = B->reg_ptr[n_parts], NE = B->enr_ptr[n_parts];
    B->n_nodes = NR;
    B->n_all = NR + NE;
    B->n_elem = B->elem_ptr[n_parts];
    B->n_free = B->free_ptr[n_parts];
    B->n_pres = B->pres_ptr[n_parts];
    B->coords.assign((size_t)3 * B->n_all, 0.0);
    B->dof.assign((size_t)dim * B->n_all, 0);
    B->vpt_ptr.push_back(0);
    int64_t v_off = 0, g_off = 0, c_off = 0;
    for (int k = 0; k < n_parts; ++k) {
        const pgf_mesh_desc& x = d[k];
        const int64_t na = x.n_all_nodes, nr = x.n_nodes;
        // node n of part k -> batch node
        auto node_of = [&](int64_t n) { return n < nr ? B->reg_ptr[k] + n : NR + B->enr_ptr[k] + (n - nr); };
        for (int64_t n = 0; n < na; ++n) {
            const int64_t g = node_of(n);
            for (int c = 0; c < 3; ++c) B->coords[(size_t)c * B->n_all + g] = x.coords[(size_t)c * na + n];
            for (int c = 0; c < dim; ++c) {
                const int32_t dk = x.dof[(size_t)c * na + n];
                B->dof[(size_t)c * B->n_all + g] =
                    dk < x.n_free ? (int32_t)(dk + B->free_ptr[k])
                                  : (int32_t)(B->n_free + B->pres_ptr[k] + (dk - x.n_free));
            }
        }
        for (int64_t i = 0; i < 4 * x.n_elem; ++i) B->conn.push_back((int32_t)(x.conn[i] + B->reg_ptr[k]));
        B->emat.insert(B->emat.end(), x.elem_mat, x.elem_mat + x.n_elem);
        for (int64_t c = 0; c < x.n_cut; ++c) {
            B->cut_elem.push_back((int32_t)(x.cut_elem[c] + B->elem_ptr[k]));
            const int ne = x.cut_n_enrich[c];
            B->cut_ne.push_back(ne);
            for (int s = 0; s < 16; ++s) {
                const int32_t en = x.cut_enrich[c * 16 + s];
                B->cut_en.push_back(s < ne ? (int32_t)(en + B->enr_ptr[k]) : en);
            }
            B->vpt_ptr.push_back(x.vpt_ptr[c + 1] + v_off);
        }
        const int64_t nv = x.n_cut ? x.vpt_ptr[x.n_cut] : 0;
        pgfh_info info;
        pgfh_problem_info(parts[k], &info);
        const int64_t ng = info.n_vgrad;
        for (int64_t q = 0; q < nv; ++q) {
            B->vpt_w.push_back(x.vpt_w[q]);
            B->vpt_mat.push_back(x.vpt_mat[q]);
            B->vpt_goff.push_back(x.vpt_grad_off[q] + g_off);
        }
        B->vpt_grad.insert(B->vpt_grad.end(), x.vpt_grad, x.vpt_grad + ng);
        B->pres.insert(B->pres.end(), pres[k].begin(), pres[k].end());
        v_off += nv;
        g_off += ng;
        c_off += x.n_cut;
    }
    B->n_cut = c_off;
    B->cpt_ptr.assign((size_t)c_off + 1, 0);
    *out = B;
    return 0;
}

PGF_API void pgfh_batch_destroy(pgfh_batch* b) { delete b; }

PGF_API int pgfh_batch_desc(pgfh_batch* b, pgf_mesh_desc* d) {
    if (!b || !d) return PGF_EARG;
    std::memset(d, 0, sizeof(*d));
    d->dim = b->dim;
    d->elem_kind = b->elem_kind;
    d->plane_strain = b->plane_strain;
    d->n_nodes = b->n_nodes;
    d->n_all_nodes = b->n_all;
    d->coords = b->coords.data();
    d->n_elem = b->n_elem;
    d->conn = b->conn.data();
    d->elem_mat = b->emat.data();
    d->n_mat = (int32_t)b->mat_E.size();
    d->mat_E = b->mat_E.data();
    d->mat_nu = b->mat_nu.data();
    d->n_cut = b->n_cut;
    d->cut_elem = b->cut_elem.data();
    d->cut_n_enrich = b->cut_ne.data();
    d->cut_enrich = b->cut_en.data();
    d->vpt_ptr = b->vpt_ptr.data();
    d->vpt_w = b->vpt_w.data();
    d->vpt_mat = b->vpt_mat.data();
    d->vpt_grad_off = b->vpt_goff.data();
    d->vpt_grad = b->vpt_grad.data();
    d->cpt_ptr = b->cpt_ptr.data();
    d->n_free = b->n_free;
    d->n_pres = b->n_pres;
    d->dof = b->dof.data();
    return 0;
}

PGF_API int pgfh_batch_offsets(pgfh_batch* b, int64_t* free_ptr, int64_t* pres_ptr, int64_t* elem_ptr,
                               int64_t* reg_ptr, int64_t* enr_ptr) {
    if (!b) return PGF_EARG;
    const size_t n = b->free_ptr.size();
    if (free_ptr) std::memcpy(free_ptr, b->free_ptr.data(), n * sizeof(int64_t));
    if (pres_ptr) std::memcpy(pres_ptr, b->pres_ptr.data(), n * sizeof(int64_t));
    if (elem_ptr) std::memcpy(elem_ptr, b->elem_ptr.data(), n * sizeof(int64_t));
    if (reg_ptr) std::memcpy(reg_ptr, b->reg_ptr.data(), n * sizeof(int64_t));
    if (enr_ptr) std::memcpy(enr_ptr, b->enr_ptr.data(), n * sizeof(int64_t));
    return 0;
}

PGF_API int pgfh_batch_get_dofs(pgfh_batch* b, int32_t* dof, double* pres) {
    if (!b) return PGF_EARG;
    if (dof) std::memcpy(dof, b->dof.data(), b->dof.size() * sizeof(int32_t));
    if (pres) std::memcpy(pres, b->pres.data(), b->pres.size() * sizeof(double));
    return 0;
}

}  // extern "C"
