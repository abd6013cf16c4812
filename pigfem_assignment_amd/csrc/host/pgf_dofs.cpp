#include "pgf_dofs.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <sstream>

namespace pgf {

namespace {
// phase timing of the set-up, printed when PGF_TIMING is set
struct PhaseTimer {
    bool on = std::getenv("PGF_TIMING") != nullptr;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    void mark(const char* what) {
        if (!on) return;
        auto n = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[pattern] %-24s %.3f s\n", what, std::chrono::duration<double>(n - t).count());
        t = n;
    }
};
}  // namespace

void DofMap::build(const Mesh& mesh, const std::vector<Dirichlet>& bcs) {
    dim = mesh.dim();
    n_all = mesh.n_all_nodes();
    // mark prescribed (node, dof); BoundaryObject keeps one value per (node, dof),
    // a later set overwriting an earlier one (src/BoundaryObject.cpp:174-186)
    std::vector<uint8_t> is_p((size_t)n_all * dim, 0);
    std::vector<double> val((size_t)n_all * dim, 0.0);
    for (const auto& b : bcs) {
        if (b.node < 0 || b.node >= mesh.n_nodes() || b.dof < 0 || b.dof >= dim)
            throw Error(PGF_ERR_ARG, "Dirichlet condition on an invalid node/dof");
        is_p[(size_t)b.node * dim + b.dof] = 1;
        val[(size_t)b.node * dim + b.dof] = b.value;
    }
    dof.assign((size_t)dim * n_all, -1);
    // periodic mesh: only the primary node of a periodic set is numbered (and
    // only its Dirichlet conditions count, as DofObject checks
    // has_dirichlet_bc for the owned primary node); the other nodes of the set
    // share its dofs (DofObject::set_periodic_nodes, src/DofObject.cpp:478-495)
    const std::vector<int32_t>* prim = mesh.periodic() ? &mesh.periodic_primary() : nullptr;
    auto numbered = [&](int64_t n) { return !prim || n >= mesh.n_nodes() || (*prim)[n] == n; };
    n_free = 0;
    n_pres = 0;
    for (int64_t n = 0; n < n_all; ++n)
        if (numbered(n))
            for (int d = 0; d < dim; ++d)
                if (is_p[(size_t)n * dim + d]) ++n_pres; else ++n_free;
    int64_t f = 0, p = 0;
    pres_value.assign(n_pres, 0.0);
    for (int64_t n = 0; n < n_all; ++n) {
        if (!numbered(n)) continue;
        for (int d = 0; d < dim; ++d) {
            if (is_p[(size_t)n * dim + d]) {
                pres_value[p] = val[(size_t)n * dim + d];
                dof[(size_t)d * n_all + n] = (int32_t)(n_free + p++);
            } else {
                dof[(size_t)d * n_all + n] = (int32_t)f++;
            }
        }
    }
    if (prim)
        for (int64_t n = 0; n < mesh.n_nodes(); ++n)
            if ((*prim)[n] != n)
                for (int d = 0; d < dim; ++d) dof[(size_t)d * n_all + n] = dof[(size_t)d * n_all + (*prim)[n]];
    if (n_free + n_pres > INT32_MAX) throw Error(PGF_ERR_LIMIT, "too many dofs for int32 ids");
}

void MeshView::finalize() {
    cut_index.assign(n_elem, -1);
    for (int64_t c = 0; c < n_cut; ++c) {
        if (cut_elem[c] < 0 || cut_elem[c] >= n_elem) throw Error(PGF_ERR_ARG, "cut element id out of range");
        if (cut_n_enrich[c] < 0 || cut_n_enrich[c] > MAX_EL_NODES - 4) throw Error(PGF_ERR_ARG, "bad enrichment count");
        cut_index[cut_elem[c]] = (int32_t)c;
    }
}

MeshView MeshView::of(const Mesh& m, std::vector<int32_t>& ce, std::vector<int32_t>& cne, std::vector<int32_t>& cen) {
    MeshView v;
    v.dim = m.dim();
    v.n_nodes = m.n_nodes();
    v.n_all = m.n_all_nodes();
    v.n_elem = m.n_elem();
    v.n_cut = (int64_t)m.cut().size();
    ce.resize(v.n_cut);
    cne.resize(v.n_cut);
    cen.assign(v.n_cut * 16, 0);
    for (int64_t c = 0; c < v.n_cut; ++c) {
        const CutElem& x = m.cut()[c];
        ce[c] = x.elem;
        cne[c] = x.n_enrich;
        for (int s = 0; s < x.n_enrich; ++s) cen[c * 16 + s] = x.enrich[s];
    }
    v.conn = m.conn().data();
    v.cut_elem = ce.data();
    v.cut_n_enrich = cne.data();
    v.cut_enrich = cen.data();
    v.finalize();
    return v;
}

void Pattern::build(const MeshView& mesh, const int32_t* dofarr, int64_t n_free_, int64_t n_pres_) {
    const int dim = mesh.dim;
    const int64_t na = mesh.n_all, ne = mesh.n_elem;
    const int32_t* conn = mesh.conn;
    PhaseTimer tm;
    // Nodes that share their dofs (the images of a periodic set) are one node
    // for the pattern: canon[n] is the first node holding n's dof ids, its rows
    // take the contributions of every element of every image
    // (Solver::combine_periodic_contributions, src/Solver.cpp:832-847).  Without
    // shared dofs canon[n] == n.
    std::vector<int32_t> canon(na);
    {
        std::vector<int32_t> first((size_t)n_free_ + n_pres_, -1);
        for (int64_t n = 0; n < na; ++n) {
            const int32_t g = dofarr[n];  // component 0
            if (g < 0 || g >= n_free_ + n_pres_) throw Error(PGF_ERR_ARG, "dof id out of range");
            if (first[g] < 0) first[g] = (int32_t)n;
            canon[n] = first[g];
        }
        for (int64_t n = 0; n < na; ++n)
            for (int d = 1; d < dim; ++d)
                if (dofarr[(size_t)d * na + n] != dofarr[(size_t)d * na + canon[n]])
                    throw Error(PGF_ERR_ARG, "nodes sharing a dof must share all their dofs");
    }
    shared_dofs = false;
    for (int64_t n = 0; n < na; ++n) shared_dofs = shared_dofs || canon[n] != n;
    // node -> element incidence over every element slot (regular and enrichment),
    // listed under the canonical node, in element order
    std::vector<int64_t> nptr(na + 1, 0);
    for (int64_t e = 0; e < ne; ++e)
        for (int k = 0; k < 4; ++k) ++nptr[canon[conn[e * 4 + k]] + 1];
    for (int64_t c = 0; c < mesh.n_cut; ++c)
        for (int s = 0; s < mesh.cut_n_enrich[c]; ++s) ++nptr[canon[mesh.n_nodes + mesh.cut_enrich[c * 16 + s]] + 1];
    for (int64_t n = 0; n < na; ++n) nptr[n + 1] += nptr[n];
    std::vector<int32_t> nel(nptr[na]);
    {
        std::vector<int64_t> fill(nptr.begin(), nptr.end() - 1);
        for (int64_t e = 0; e < ne; ++e)
            for (int k = 0; k < 4; ++k) nel[fill[canon[conn[e * 4 + k]]]++] = (int32_t)e;
        for (int64_t c = 0; c < mesh.n_cut; ++c)
            for (int s = 0; s < mesh.cut_n_enrich[c]; ++s)
                nel[fill[canon[mesh.n_nodes + mesh.cut_enrich[c * 16 + s]]]++] = mesh.cut_elem[c];
        if (shared_dofs)
#pragma omp parallel for schedule(dynamic, 4096)
            for (int64_t n = 0; n < na; ++n) std::sort(nel.begin() + nptr[n], nel.begin() + nptr[n + 1]);
    }
    tm.mark("incidence");
    // node neighbour lists (sorted canonical node ids)
    std::vector<int64_t> nbptr(na + 1, 0);
    std::vector<std::vector<int32_t>> nbtmp(na);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t n = 0; n < na; ++n) {
        auto& v = nbtmp[n];
        for (int64_t k = nptr[n]; k < nptr[n + 1]; ++k) {
            int64_t e = nel[k];
            int nt = mesh.n_total(e);
            for (int s = 0; s < nt; ++s) v.push_back(canon[mesh.node(e, s)]);
        }
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    for (int64_t n = 0; n < na; ++n) nbptr[n + 1] = nbptr[n] + (int64_t)nbtmp[n].size();
    std::vector<int32_t> nb(nbptr[na]);
    std::vector<uint8_t> off_free(nbptr[na]), off_pres(nbptr[na]);
    std::vector<int32_t> rowlen(na), nfree_row(na);
    const int32_t nf = (int32_t)n_free_;
    auto is_free = [&](int64_t node, int d) { return dofarr[(size_t)d * na + node] < nf; };
    bool overflow = false;
#pragma omp parallel for schedule(dynamic, 4096) reduction(|| : overflow)
    for (int64_t n = 0; n < na; ++n) {
        const auto& v = nbtmp[n];
        int fr = 0;
        for (size_t i = 0; i < v.size(); ++i)
            for (int d = 0; d < dim; ++d) fr += is_free(v[i], d);
        int cf = 0, cp = fr;
        for (size_t i = 0; i < v.size(); ++i) {
            nb[nbptr[n] + i] = v[i];
            if (cf > 255 || cp > 255) overflow = true;
            off_free[nbptr[n] + i] = (uint8_t)cf;
            off_pres[nbptr[n] + i] = (uint8_t)cp;
            for (int d = 0; d < dim; ++d) { if (is_free(v[i], d)) ++cf; else ++cp; }
        }
        rowlen[n] = (int32_t)(v.size() * dim);
        nfree_row[n] = fr;
    }
    if (overflow) throw Error(PGF_ERR_LIMIT, "row wider than 255 columns");
    tm.mark("neighbour lists");
    nbtmp.clear();
    nbtmp.shrink_to_fit();
    // row layout: every row of node n = free dofs of its neighbours (node order),
    // then their prescribed dofs
    n_rows = n_free_ + n_pres_;
    std::vector<int64_t> row_node(n_rows), row_len(n_rows);
    for (int64_t n = 0; n < na; ++n)
        for (int d = 0; d < dim; ++d) {
            int32_t r = dofarr[(size_t)d * na + n];
            if (r < 0 || r >= n_rows) throw Error(PGF_ERR_ARG, "dof id out of range");
            row_node[r] = canon[n];
        }
    rowptr.assign(n_rows + 1, 0);
    for (int64_t r = 0; r < n_rows; ++r) rowptr[r + 1] = rowptr[r] + rowlen[row_node[r]];
    nnz = rowptr[n_rows];
    colidx.resize(nnz);
    rowsplit.resize(n_rows);
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t r = 0; r < n_rows; ++r) {
        int64_t n = row_node[r];
        int64_t p = rowptr[r];
        int64_t pf = p, pp = p + nfree_row[n];
        for (int64_t i = nbptr[n]; i < nbptr[n + 1]; ++i)
            for (int d = 0; d < dim; ++d) {
                int32_t c = dofarr[(size_t)d * na + nb[i]];
                if (c < nf) colidx[pf++] = c; else colidx[pp++] = c;
            }
        rowsplit[r] = nfree_row[n];
    }
    tm.mark("csr rows");
    // element scatter offsets
    auto lookup = [&](int32_t a, int32_t b, uint8_t& of, uint8_t& op) {
        a = canon[a];
        b = canon[b];
        const int32_t* beg = &nb[nbptr[a]];
        const int32_t* end = &nb[nbptr[a + 1]];
        const int32_t* it = std::lower_bound(beg, end, b);
        int64_t k = nbptr[a] + (it - beg);
        of = off_free[k];
        op = off_pres[k];
    };
    elem_off.assign((size_t)ne * 32, 0);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                lookup(conn[e * 4 + i], conn[e * 4 + j], elem_off[(size_t)e * 32 + (i * 4 + j) * 2],
                       elem_off[(size_t)e * 32 + (i * 4 + j) * 2 + 1]);
    cut_off_ptr.assign(mesh.n_cut + 1, 0);
    for (int64_t c = 0; c < mesh.n_cut; ++c) {
        int nt = 4 + mesh.cut_n_enrich[c];
        cut_off_ptr[c + 1] = cut_off_ptr[c] + 2 * nt * nt;
    }
    cut_off.assign(cut_off_ptr.back(), 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < mesh.n_cut; ++c) {
        int nt = 4 + mesh.cut_n_enrich[c];
        int64_t e = mesh.cut_elem[c];
        for (int i = 0; i < nt; ++i)
            for (int j = 0; j < nt; ++j)
                lookup(mesh.node(e, i), mesh.node(e, j), cut_off[cut_off_ptr[c] + (i * nt + j) * 2],
                       cut_off[cut_off_ptr[c] + (i * nt + j) * 2 + 1]);
    }
    tm.mark("scatter offsets");
    // every scatter slot must land on its own column inside its own row: this
    // check keeps the device scatter in bounds by construction
    bool bad = false;
    auto check_elem = [&](int nt, auto node_of, const uint8_t* off) {
        for (int i = 0; i < nt; ++i) {
            const int32_t a = node_of(i);
            for (int di = 0; di < dim; ++di) {
                const int32_t r = dofarr[(size_t)di * na + a];
                const int64_t base = rowptr[r], end = rowptr[r + 1];
                for (int j = 0; j < nt; ++j) {
                    const int32_t b = node_of(j);
                    int cf = 0, cp = 0;
                    for (int dj = 0; dj < dim; ++dj) {
                        const int32_t c = dofarr[(size_t)dj * na + b];
                        const int64_t pos = c < nf ? base + off[(i * nt + j) * 2] + cf++ : base + off[(i * nt + j) * 2 + 1] + cp++;
                        if (pos < base || pos >= end || colidx[pos] != c) return true;
                    }
                }
            }
        }
        return false;
    };
#pragma omp parallel for schedule(static) reduction(|| : bad)
    for (int64_t e = 0; e < ne; ++e)
        if (check_elem(4, [&](int s) { return conn[e * 4 + s]; }, &elem_off[(size_t)e * 32])) bad = true;
#pragma omp parallel for schedule(dynamic, 64) reduction(|| : bad)
    for (int64_t c = 0; c < mesh.n_cut; ++c) {
        const int64_t e = mesh.cut_elem[c];
        if (check_elem(4 + mesh.cut_n_enrich[c], [&](int s) { return mesh.node(e, s); }, &cut_off[cut_off_ptr[c]]))
            bad = true;
    }
    if (bad) throw Error(PGF_ERR_STATE, "scatter offsets do not match the CSR pattern");
    tm.mark("offset validation");
    // row-gather plan from the node -> element incidence (element order per node)
    inc_ptr.assign(nptr.begin(), nptr.end());
    inc_elem.assign(nel.begin(), nel.end());
    inc_slot.assign(nel.size(), 0);
    inc_off.assign(nel.size(), 0);
    bool two_images = false;
#pragma omp parallel for schedule(dynamic, 4096) reduction(|| : two_images)
    for (int64_t n = 0; n < na; ++n)
        for (int64_t k = nptr[n]; k < nptr[n + 1]; ++k) {
            const int64_t e = nel[k];
            const int nt = mesh.n_total(e);
            int la = -1;
            for (int s = 0; s < nt; ++s)
                if (canon[mesh.node(e, s)] == n) {
                    if (la >= 0) two_images = true;
                    if (la < 0) la = s;
                }
            inc_slot[k] = (uint8_t)la;
            if (mesh.cut_index[e] < 0) {
                uint64_t w = 0;
                for (int b = 0; b < 4; ++b) {
                    w |= (uint64_t)elem_off[(size_t)e * 32 + (la * 4 + b) * 2] << (16 * b);
                    w |= (uint64_t)elem_off[(size_t)e * 32 + (la * 4 + b) * 2 + 1] << (16 * b + 8);
                }
                inc_off[k] = w;
            }
        }
    if (two_images) throw Error(PGF_ERR_ARG, "an element holds two nodes that share their dofs");
    tm.mark("gather plan");
    // greedy colouring in element order over the regular nodes (elements that
    // share an enrichment node always share that edge's regular nodes)
    color.assign(ne, 0);
    std::vector<uint64_t> mask(mesh.n_nodes, 0);
    n_colors = 0;
    for (int64_t e = 0; e < ne; ++e) {
        uint64_t used = 0;
        for (int k = 0; k < 4; ++k) used |= mask[canon[conn[e * 4 + k]]];
        if (used == ~0ull) throw Error(PGF_ERR_LIMIT, "more than 64 colours needed");
        int c = __builtin_ctzll(~used);
        color[e] = c;
        for (int k = 0; k < 4; ++k) mask[canon[conn[e * 4 + k]]] |= 1ull << c;
        n_colors = std::max(n_colors, c + 1);
    }
    tm.mark("colouring");
}

}  // namespace pgf
