// Host mesh: generation, IGFEM detection, enrichment and cut-element tables.
// Reference correspondence is given per routine; the detection case tables are
// the reference's (src/quad4.cpp:188-560, src/tet4.cpp:228-520) because the
// sub-element node order fixes quadrature-point order, the internal-variable
// layout and the cohesive orientation that the damage flags are reported in.
#include "pgf_mesh.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <sstream>

#include "pgf_geom.hpp"
#include "pgf_igfem.hpp"

namespace pgf {

namespace {

[[noreturn]] void fail(int code, const std::string& msg) { throw Error(code, msg); }

}  // namespace

// ---------------------------------------------------------------------------
// Mesh generation (src/Mesh.cpp:191-583): nodes x fastest, then y, then z;
// QUAD4 counter-clockwise; each cube split into 6 tets sharing the 0-6 diagonal.
void Mesh::generate_mesh(ElemKind kind, const double* mins, const double* maxs, const int* N) {
    _kind = kind;
    _dim = kind == QUAD4 ? 2 : 3;
    _nn = 4;
    for (int d = 0; d < 3; ++d) {
        _N[d] = d < _dim ? N[d] : 0;
        _min[d] = d < _dim ? mins[d] : 0.0;
        _max[d] = d < _dim ? maxs[d] : 0.0;
    }
    const int64_t Nx = _N[0], Ny = _N[1], Nz = _N[2];
    if (Nx <= 0 || Ny <= 0 || (_dim == 3 && Nz <= 0)) fail(PGF_ERR_ARG, "mesh divisions must be positive");
    double hx = (_max[0] - _min[0]) / Nx, hy = (_max[1] - _min[1]) / Ny;
    double hz = _dim == 3 ? (_max[2] - _min[2]) / Nz : 0.0;
    _n_nodes = (Nx + 1) * (Ny + 1) * (Nz + 1);
    for (int d = 0; d < 3; ++d) _X[d].assign(_n_nodes, 0.0);
    int64_t id = 0;
    for (int64_t z = 0; z <= Nz; ++z)
        for (int64_t y = 0; y <= Ny; ++y)
            for (int64_t x = 0; x <= Nx; ++x, ++id) {
                _X[0][id] = _min[0] + x * hx;
                _X[1][id] = _min[1] + y * hy;
                _X[2][id] = _min[2] + z * hz;
            }
    if (_dim == 2) {
        _n_elem = Nx * Ny;
        _conn.resize(_n_elem * 4);
        int64_t e = 0;
        for (int64_t y = 0; y < Ny; ++y)
            for (int64_t x = 0; x < Nx; ++x, ++e) {
                int32_t* c = &_conn[e * 4];
                c[0] = (int32_t)(x + y * (Nx + 1));
                c[1] = (int32_t)(x + 1 + y * (Nx + 1));
                c[2] = (int32_t)(x + 1 + (y + 1) * (Nx + 1));
                c[3] = (int32_t)(x + (y + 1) * (Nx + 1));
            }
    } else {
        _n_elem = 6 * Nx * Ny * Nz;
        if (_n_elem > INT32_MAX || _n_nodes > INT32_MAX) fail(PGF_ERR_LIMIT, "mesh too large for int32 ids");
        _conn.resize(_n_elem * 4);
        int64_t e = 0;
        const int64_t sy = Nx + 1, sz = (Nx + 1) * (Ny + 1);
        for (int64_t z = 0; z < Nz; ++z)
            for (int64_t y = 0; y < Ny; ++y)
                for (int64_t x = 0; x < Nx; ++x) {
                    int32_t v[8];
                    v[0] = (int32_t)(x + y * sy + z * sz);
                    v[1] = (int32_t)(x + 1 + y * sy + z * sz);
                    v[2] = (int32_t)(x + 1 + (y + 1) * sy + z * sz);
                    v[3] = (int32_t)(x + (y + 1) * sy + z * sz);
                    v[4] = (int32_t)(x + y * sy + (z + 1) * sz);
                    v[5] = (int32_t)(x + 1 + y * sy + (z + 1) * sz);
                    v[6] = (int32_t)(x + 1 + (y + 1) * sy + (z + 1) * sz);
                    v[7] = (int32_t)(x + (y + 1) * sy + (z + 1) * sz);
                    static const int T[6][4] = {{0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6},
                                                {0, 5, 6, 4}, {1, 5, 6, 0}, {1, 6, 2, 0}};
                    for (int t = 0; t < 6; ++t, ++e)
                        for (int k = 0; k < 4; ++k) _conn[e * 4 + k] = v[T[t][k]];
                }
    }
    _elem_mat.assign(_n_elem, 0);
    // lowest element id holding each node (src/Mesh.cpp:2379-2404 appends in
    // element order, so _node_elem[n][0] is the lowest id)
    _first_elem.assign(_n_nodes, -1);
    for (int64_t e = 0; e < _n_elem; ++e)
        for (int k = 0; k < 4; ++k) {
            int32_t n = _conn[e * 4 + k];
            if (_first_elem[n] < 0) _first_elem[n] = (int32_t)e;
        }
    _cut.clear();
    _cut_index.assign(_n_elem, -1);
    _enrich_edge.clear();
    _nodes_detected = _elements_detected = _enriched = false;
}

std::vector<int32_t> Mesh::nodeset(const std::string& name) const {
    std::vector<int32_t> out;
    const int64_t Nx = _N[0], Ny = _N[1], Nz = _N[2];
    int64_t id = 0;
    for (int64_t z = 0; z <= Nz; ++z)
        for (int64_t y = 0; y <= Ny; ++y)
            for (int64_t x = 0; x <= Nx; ++x, ++id) {
                bool in = false;
                if (name == "LEFT") in = x == 0;
                else if (name == "RIGHT") in = x == Nx;
                else if (_dim == 2 && name == "BOTTOM") in = y == 0;
                else if (_dim == 2 && name == "TOP") in = y == Ny;
                else if (_dim == 3 && name == "FRONT") in = y == 0;
                else if (_dim == 3 && name == "BACK") in = y == Ny;
                else if (_dim == 3 && name == "BOTTOM") in = z == 0;
                else if (_dim == 3 && name == "TOP") in = z == Nz;
                if (in) out.push_back((int32_t)id);
            }
    if (out.empty()) fail(PGF_ERR_ARG, "unknown nodeset " + name);
    return out;
}

void Mesh::build_periodic_nodesets() {
    if (_n_nodes == 0) fail(PGF_ERR_STATE, "generate the mesh before its periodic sets");
    const int64_t Nx = _N[0], Ny = _N[1], Nz = _N[2];
    if (Nx < 2 || Ny < 2 || (_dim == 3 && Nz < 2)) fail(PGF_ERR_ARG, "a periodic mesh needs at least 2 divisions per side");
    _periodic_primary.resize(_n_nodes);
    int64_t id = 0;
    for (int64_t z = 0; z <= Nz; ++z)
        for (int64_t y = 0; y <= Ny; ++y)
            for (int64_t x = 0; x <= Nx; ++x, ++id) {
                const int64_t px = x % Nx, py = y % Ny, pz = _dim == 3 ? z % Nz : 0;
                _periodic_primary[id] = (int32_t)(px + py * (Nx + 1) + pz * (Nx + 1) * (Ny + 1));
            }
}

// ---------------------------------------------------------------------------
// Node detection (src/Mesh.cpp:2838-2891, src/InclusionCircle.cpp:150-163,
// src/InclusionSphere.cpp:163-178): inclusions scanned last-to-first, a node
// within 1e-12 (relative) of a surface is pushed 5% of an element edge along
// the radial direction and the scan restarts.
void Mesh::move_node(int64_t n, const Inclusion& inc) {
    double nrm[3] = {_X[0][n] - inc.c[0], _X[1][n] - inc.c[1], inc.type == 1 ? _X[2][n] - inc.c[2] : 0.0};
    double mag = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    for (int d = 0; d < 3; ++d) nrm[d] /= mag;
    int64_t e = _first_elem[n];
    int64_t other = _conn[e * 4 + 0];
    if (other == n) other = _conn[e * 4 + 1];
    double dx = _X[0][n] - _X[0][other], dy = _X[1][n] - _X[1][other], dz = _X[2][n] - _X[2][other];
    double move = 0.05 * std::sqrt(dx * dx + dy * dy + dz * dz);
    for (int d = 0; d < 3; ++d) _X[d][n] = _X[d][n] + move * nrm[d];
    ++_n_moved;
}

int Mesh::detect_node(int64_t n, int& times_moved) {
    for (;;) {
        const int r = igfem::scan_node(_incl.data(), (int)_incl.size(), _opt.sphere_zlim_as_reference, _X[0][n],
                                       _X[1][n], _X[2][n]);
        if (r >= -1) return r;
        if (times_moved >= 5) fail(PGF_ERR_GEOMETRY, "unable to move node off an inclusion surface");
        move_node(n, _incl[-2 - r]);
        ++times_moved;
    }
}

void Mesh::analyze_nodes() {
    _node_detect.assign(_n_nodes, -1);
    _n_moved = 0;
    if (_incl.empty()) { _nodes_detected = true; return; }
    for (auto& inc : _incl)
        if ((inc.type == 0) != (_dim == 2)) fail(PGF_ERR_ARG, "inclusion type does not match mesh dimension");
    for (int64_t n = 0; n < _n_nodes; ++n) {
        int moved = 0;
        _node_detect[n] = detect_node(n, moved);
    }
    _nodes_detected = true;
}

// ---------------------------------------------------------------------------
// Element detection.  Only the cases a single inclusion can produce (one
// interface per element) are tabulated; elements touched by two inclusions are
// rejected with PGF_ERR_UNSUPPORTED (the reference rejects them for tets,
// src/tet4.cpp:259-263, and for quads with opposite-corner patterns).
void Mesh::edge_nodes(int edge, int& a, int& b) const { igfem::edge_nodes(_kind, edge, a, b); }

// src/quad4.cpp:188-560 and src/tet4.cpp:228-520 (tables in pgf_igfem.hpp)
void Mesh::detect_element(int64_t e, const int* nd, CutElem& ce, bool& cut) {
    const int st = _kind == QUAD4 ? igfem::classify_quad4(nd, _elem_mat[e], _incl.data(), _opt.cohesive, ce, cut)
                                  : igfem::classify_tet4(nd, _elem_mat[e], _incl.data(), _opt.cohesive, ce, cut);
    if (st == igfem::CLS_OK) return;
    std::ostringstream s;
    const bool quad = _kind == QUAD4;
    if (st == igfem::CLS_MULTI)
        s << "element " << e << " is cut by more than one " << (quad ? "inclusion" : "interface") << " (unsupported)";
    else if (st == igfem::CLS_ONE_UNKNOWN)
        s << (quad ? "quadrilateral" : "tetrahedron") << " has 1 node separated in an unknown manner";
    else
        s << (quad ? "quadrilateral element " : "tetrahedron ") << e
          << " has 2 nodes separated in an unknown manner (refine)";
    fail(PGF_ERR_UNSUPPORTED, s.str());
}

// src/Mesh.cpp:2922-3027
void Mesh::analyze_elements() {
    if (!_nodes_detected) fail(PGF_ERR_STATE, "analyze_nodes must run before analyze_elements");
    _cut.clear();
    _cut_index.assign(_n_elem, -1);
    if (_incl.empty()) { _elements_detected = true; return; }
    for (int64_t e = 0; e < _n_elem; ++e) {
        int nd[4];
        for (int k = 0; k < 4; ++k) nd[k] = _node_detect[_conn[e * 4 + k]];
        CutElem ce;
        bool cut = false;
        detect_element(e, nd, ce, cut);
        if (!cut) {
            if (nd[0] != -1) _elem_mat[e] = _incl[nd[0]].mat;
            continue;
        }
        ce.elem = (int32_t)e;
        igfem::cohesive_laws(ce, _incl.data(), _opt.default_coh_law);
        _cut_index[e] = (int32_t)_cut.size();
        _cut.push_back(ce);
    }
    _elements_detected = true;
}

// src/Mesh.cpp:3064-3202 (serial branch) and :3600-3656: one enrichment node per
// distinct (sorted edge, inclusion) pair, numbered in that sorted order, then
// one mirror node per enrichment node for cohesive meshes.
void Mesh::add_enrichments() {
    if (!_elements_detected) fail(PGF_ERR_STATE, "analyze_elements must run before add_enrichments");
    std::vector<std::array<int32_t, 3>> keys;
    keys.reserve(_cut.size() * 4);
    for (auto& ce : _cut)
        for (int i = 0; i < ce.n_cut; ++i) {
            int a, b;
            edge_nodes(ce.cut_edge[i], a, b);
            int32_t g0 = _conn[(int64_t)ce.elem * 4 + a], g1 = _conn[(int64_t)ce.elem * 4 + b];
            keys.push_back({std::min(g0, g1), std::max(g0, g1), ce.enrich_incl[i]});
        }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const int64_t n_init = (int64_t)keys.size();
    const int64_t n_total = _opt.cohesive ? 2 * n_init : n_init;
    _enrich_edge = keys;
    if (_opt.cohesive) _enrich_edge.insert(_enrich_edge.end(), keys.begin(), keys.end());
    for (int d = 0; d < 3; ++d) _X[d].resize(_n_nodes + n_total);
    for (int64_t k = 0; k < n_init; ++k) {
        double p0[3], p1[3], x[3];
        for (int d = 0; d < 3; ++d) { p0[d] = _X[d][keys[k][0]]; p1[d] = _X[d][keys[k][1]]; }
        if (!igfem::enrichment_point(_incl[keys[k][2]], p0, p1, x))
            fail(PGF_ERR_GEOMETRY, "no valid intersection on cut edge");
        for (int d = 0; d < 3; ++d) {
            _X[d][_n_nodes + k] = x[d];
            if (_opt.cohesive) _X[d][_n_nodes + n_init + k] = x[d];
        }
    }
    for (auto& ce : _cut)
        for (int i = 0; i < ce.n_cut; ++i) {
            int a, b;
            edge_nodes(ce.cut_edge[i], a, b);
            int32_t g0 = _conn[(int64_t)ce.elem * 4 + a], g1 = _conn[(int64_t)ce.elem * 4 + b];
            std::array<int32_t, 3> key = {std::min(g0, g1), std::max(g0, g1), ce.enrich_incl[i]};
            auto it = std::lower_bound(keys.begin(), keys.end(), key);
            int32_t k = (int32_t)(it - keys.begin());
            ce.enrich[i] = k;
            if (_opt.cohesive) ce.enrich[i + ce.n_cut] = (int32_t)(k + n_init);
        }
    _enriched = true;
}

int32_t Mesh::elem_node(int64_t e, int slot) const {
    if (slot < 4) return _conn[e * 4 + slot];
    const CutElem& ce = _cut[_cut_index[e]];
    return (int32_t)(_n_nodes + ce.enrich[slot - 4]);
}

int Mesh::elem_n_total(int64_t e) const {
    int ci = _cut_index.empty() ? -1 : _cut_index[e];
    return ci < 0 ? 4 : 4 + _cut[ci].n_enrich;
}

// ---------------------------------------------------------------------------
// Cut-element integration tables (src/Mesh.cpp:1495-1602).  For every
// integration sub-element quadrature point: child shape functions and child
// Jacobian J, global point, inverse map into the parent, parent gradients for
// the regular nodes and child gradients for the enrichment nodes the child
// contains; weight J*W.  In 3D every sub-element is a linear tet whose
// gradients are the same at all five points, so its points are merged into one
// with the summed weight.  Cohesive points: shape values, rotation matrix
// (tangents, then unit normal) and dA*W (src/CohesiveElem.cpp:130-242).
void Mesh::store_shape_functions() {
    if (!_enriched && !_cut.empty()) fail(PGF_ERR_STATE, "add_enrichments must run before store_shape_functions");
    const int dim = _dim;
    const int64_t nc = (int64_t)_cut.size();
    // sizes per cut element, then offsets, then every element's entries in parallel
    vpt_ptr.assign(nc + 1, 0);
    cpt_ptr.assign(nc + 1, 0);
    std::vector<int64_t> gptr(nc + 1, 0);
    for (int64_t i = 0; i < nc; ++i) {
        int nv, ncp;
        igfem::table_sizes(dim, _cut[i], nv, ncp);
        vpt_ptr[i + 1] = vpt_ptr[i] + nv;
        cpt_ptr[i + 1] = cpt_ptr[i] + ncp;
        gptr[i + 1] = gptr[i] + (int64_t)nv * (4 + _cut[i].n_enrich) * dim;
    }
    const int64_t NV = vpt_ptr[nc], NC = cpt_ptr[nc];
    vpt_w.assign(NV, 0.0); vpt_mat.assign(NV, 0); vpt_grad_off.assign(NV, 0); vpt_grad.assign(gptr[nc], 0.0);
    cpt_w.assign(NC, 0.0); cpt_N.assign(4 * NC, 0.0); cpt_R.assign(9 * NC, 0.0); cpt_nodes.assign(8 * NC, 0);
    cpt_nh.assign(NC, 0); cpt_law.assign(NC, 0);
    const geom::PolygonTable& poly = geom::polygon_table();
    int64_t bad = nc;
    int bad_code = 0;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < nc; ++i) {
        const CutElem& ce = _cut[i];
        const int nt = 4 + ce.n_enrich;
        double X[MAX_EL_NODES][3];
        for (int s = 0; s < nt; ++s) {
            int32_t g = elem_node(ce.elem, s);
            for (int d = 0; d < 3; ++d) X[s][d] = _X[d][g];
        }
        igfem::TableOut o;
        o.vw = &vpt_w[vpt_ptr[i]]; o.vmat = &vpt_mat[vpt_ptr[i]]; o.vgoff = &vpt_grad_off[vpt_ptr[i]];
        o.vgrad = vpt_grad.data() + gptr[i]; o.goff0 = gptr[i];
        o.cw = cpt_w.data() + cpt_ptr[i]; o.cN = cpt_N.data() + 4 * cpt_ptr[i]; o.cR = cpt_R.data() + 9 * cpt_ptr[i];
        o.cnodes = cpt_nodes.data() + 8 * cpt_ptr[i]; o.cnh = cpt_nh.data() + cpt_ptr[i];
        o.claw = cpt_law.data() + cpt_ptr[i];
        const int st = igfem::cut_tables(dim, _kind, ce, X, _opt.tet_weight_scale, poly, o);
        if (st != igfem::TAB_OK) {
#pragma omp critical
            if (i < bad) { bad = i; bad_code = st; }
        }
    }
    if (bad < nc) {
        std::ostringstream s;
        if (bad_code == igfem::TAB_NEG_JACOBIAN) s << "negative sub-element Jacobian in element " << _cut[bad].elem;
        else if (bad_code == igfem::TAB_OUTSIDE) s << "bad inverse map detected";
        else s << "bad inverse map in assembly detected";
        fail(PGF_ERR_GEOMETRY, s.str());
    }
}

}  // namespace pgf
