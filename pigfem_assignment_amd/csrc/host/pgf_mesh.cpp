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

namespace pgf {

namespace {

[[noreturn]] void fail(int code, const std::string& msg) { throw Error(code, msg); }

// Level-set value of an inclusion along the edge p0 + a (p1 - p0), as the
// reference's int_f / int_df (src/InclusionCircle.cpp:248-287,
// src/InclusionSphere.cpp:270-326).
struct EdgeLevelSet {
    const Inclusion* inc;
    double p0[3], p1[3];
    int nd;  // 2 for circles, 3 for spheres
    double f(double a) const {
        double s = 0.0;
        for (int d = 0; d < nd; ++d) {
            double g = (p0[d] + a * (p1[d] - p0[d]) - inc->c[d]) / inc->r;
            s += g * g;
        }
        return s - 1.0;
    }
    double df(double a) const {
        double s = 0.0;
        for (int d = 0; d < nd; ++d) {
            double g = p0[d] + a * (p1[d] - p0[d]) - inc->c[d];
            s += g * (p1[d] - p0[d]);
        }
        return (1.0 / (inc->r * inc->r)) * 2.0 * s;
    }
};

// GSL 2.1 Steffenson iteration (roots/steffenson.c): Newton steps, with the
// Aitken delta-squared estimate reported from the third iterate on.
// GSL 2.1 Brent (roots/brent.c) as the bracketing fallback.  Driven exactly as
// Inclusion::find_intersection drives them (src/Inclusion.cpp:32-123):
// start 0.5, |delta| < 1e-7, 25 iterations, Brent on [0,1] when Steffenson
// runs out of iterations or leaves [0,1].
double find_root(const EdgeLevelSet& ls) {
    if (ls.f(0.0) * ls.f(1.0) >= 0) fail(PGF_ERR_GEOMETRY, "no valid intersection on cut edge");
    // Steffenson
    double x = 0.5, x1 = 0.0, fx = ls.f(x), dfx = ls.df(x), root = 0.5;
    int count = 1;
    double alpha = 0.5, alpha0;
    int iter = 0;
    const int max_iter = 25;
    bool conv = false;
    do {
        ++iter;
        if (dfx != 0.0) {
            double xn = x - fx / dfx;
            double fn = ls.f(xn), dfn = ls.df(xn);
            double xold1 = x1;
            x1 = x;
            double xprev = x;
            x = xn; fx = fn; dfx = dfn;
            if (count < 3) { root = xn; ++count; }
            else {
                double u = xprev - xold1, v = xn - 2 * xprev + xold1;
                root = (v == 0) ? xn : xold1 - u * u / v;
            }
        }
        alpha0 = alpha;
        alpha = root;
        conv = std::fabs(alpha - alpha0) < 1e-7 || alpha == alpha0;
    } while (!conv && iter < max_iter);
    if (iter == max_iter || alpha < 0.0 || alpha > 1.0) {
        // Brent
        double a = 0.0, b = 1.0, c = 1.0, fa = ls.f(0.0), fb = ls.f(1.0), fc = fb;
        double d = 1.0, e = 1.0, lo = 0.0, hi = 1.0;
        alpha = 0.5;
        for (int it = 0; it < 100; ++it) {
            bool ac_equal = false;
            if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) { ac_equal = true; c = a; fc = fa; d = b - a; e = b - a; }
            if (std::fabs(fc) < std::fabs(fb)) { ac_equal = true; a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
            double tol = 0.5 * DBL_EPSILON * std::fabs(b), m = 0.5 * (c - b);
            if (fb == 0) { alpha = b; lo = hi = b; }
            else if (std::fabs(m) <= tol) { alpha = b; lo = std::min(b, c); hi = std::max(b, c); }
            else {
                if (std::fabs(e) < tol || std::fabs(fa) <= std::fabs(fb)) { d = m; e = m; }
                else {
                    double p, q, r, s = fb / fa;
                    if (ac_equal) { p = 2 * m * s; q = 1 - s; }
                    else {
                        q = fa / fc; r = fb / fc;
                        p = s * (2 * m * q * (q - r) - (b - a) * (r - 1));
                        q = (q - 1) * (r - 1) * (s - 1);
                    }
                    if (p > 0) q = -q; else p = -p;
                    if (2 * p < std::min(3 * m * q - std::fabs(tol * q), std::fabs(e * q))) { e = d; d = p / q; }
                    else { d = m; e = m; }
                }
                a = b; fa = fb;
                if (std::fabs(d) > tol) b += d; else b += (m > 0 ? +tol : -tol);
                fb = ls.f(b);
                alpha = b;
                double cc = c;
                if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) cc = a;
                lo = std::min(b, cc); hi = std::max(b, cc);
            }
            if (std::fabs(lo - hi) < 1e-7 || lo == hi) break;
        }
    }
    return alpha;
}

// element edge -> local nodes (src/quad4.cpp:23-29, src/tet4.cpp:22-30)
const int kQuadEdge[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
const int kTetEdge[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};

void set_subs(CutElem& ce, std::initializer_list<std::initializer_list<int>> subs) {
    ce.n_sub = 0;
    for (auto& s : subs) {
        int k = 0;
        for (int v : s) ce.sub_nodes[ce.n_sub][k++] = (uint8_t)v;
        ce.sub_nn[ce.n_sub++] = (uint8_t)k;
    }
}
void set_cohs(CutElem& ce, std::initializer_list<std::initializer_list<int>> cohs) {
    ce.n_coh = 0;
    for (auto& s : cohs) {
        int k = 0;
        for (int v : s) ce.coh_nodes[ce.n_coh][k++] = (uint8_t)v;
        ce.coh_nn[ce.n_coh++] = (uint8_t)k;
    }
}
void set_edges(CutElem& ce, std::initializer_list<int> edges) {
    ce.n_cut = 0;
    for (int v : edges) ce.cut_edge[ce.n_cut++] = (uint8_t)v;
}

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
    const int ni = (int)_incl.size();
    for (int j = ni - 1; j >= 0; --j) {
        const Inclusion& inc = _incl[j];
        double x = _X[0][n], y = _X[1][n], z = _X[2][n];
        double lo[3], hi[3];
        if (inc.type == 0) {
            lo[0] = inc.c[0] - inc.r; lo[1] = inc.c[1] - inc.r; lo[2] = -DBL_MAX;
            hi[0] = inc.c[0] + inc.r; hi[1] = inc.c[1] + inc.r; hi[2] = DBL_MAX;
        } else {
            lo[0] = inc.c[0] - inc.r; lo[1] = inc.c[1] - inc.r; lo[2] = inc.c[2] - inc.r;
            hi[0] = inc.c[0] + inc.r; hi[1] = inc.c[1] + inc.r;
            hi[2] = _opt.sphere_zlim_as_reference ? inc.c[1] + inc.r : inc.c[2] + inc.r;
        }
        if (!(lo[0] <= x && lo[1] <= y && lo[2] <= z && x <= hi[0] && y <= hi[1] && z <= hi[2])) continue;
        double val;
        if (inc.type == 0) {
            double gx = x - inc.c[0], gy = y - inc.c[1];
            val = (std::sqrt(gx * gx + gy * gy) - inc.r) / inc.r;
        } else {
            double gx = (x - inc.c[0]) / inc.r, gy = (y - inc.c[1]) / inc.r, gz = (z - inc.c[2]) / inc.r;
            val = gx * gx + gy * gy + gz * gz - 1.0;
        }
        const double eps = 1e-12;
        if (val > eps) continue;
        if (val < -eps) return j;
        if (times_moved >= 5) fail(PGF_ERR_GEOMETRY, "unable to move node off an inclusion surface");
        move_node(n, inc);
        ++times_moved;
        j = ni;
    }
    return -1;
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
void Mesh::edge_nodes(int edge, int& a, int& b) const {
    if (_kind == QUAD4) { a = kQuadEdge[edge][0]; b = kQuadEdge[edge][1]; }
    else { a = kTetEdge[edge][0]; b = kTetEdge[edge][1]; }
}

static void unique_counts(const int* nd, int nn, int* vec, int& nu, int* counts) {
    int tmp[4];
    for (int i = 0; i < nn; ++i) tmp[i] = nd[i];
    std::sort(tmp, tmp + nn);
    nu = 0;
    for (int i = 0; i < nn; ++i)
        if (nu == 0 || vec[nu - 1] != tmp[i]) vec[nu++] = tmp[i];
    for (int i = 0; i < nu; ++i) counts[i] = (int)std::count(nd, nd + nn, vec[i]);
}

// src/quad4.cpp:188-560
void Mesh::detect_quad4(int64_t e, const int* nd, CutElem& ce, bool& cut) {
    const int n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
    const bool n0e1 = n0 == n1, n0e2 = n0 == n2, n0e3 = n0 == n3, n1e2 = n1 == n2, n1e3 = n1 == n3;
    int vec[4], cnt[4], nu;
    unique_counts(nd, 4, vec, nu, cnt);
    cut = nu > 1;
    if (!cut) return;
    if (nu != 2 || vec[0] >= 0) {
        std::ostringstream s;
        s << "element " << e << " is cut by more than one inclusion (unsupported)";
        fail(PGF_ERR_UNSUPPORTED, s.str());
    }
    auto M = [&](int v) { return v < 0 ? _elem_mat[e] : _incl[v].mat; };
    const bool coh = _opt.cohesive;
    int sm[2];
    if (cnt[0] == 1 || cnt[0] == 3) {
        if (n1e2 && n1e3) {          // node 0 separate
            set_edges(ce, {3, 0}); sm[0] = M(n0); sm[1] = M(n2);
            if (!coh) set_subs(ce, {{0, 5, 4}, {1, 2, 3, 4, 5}});
            else if (n0 < n1) { set_subs(ce, {{0, 5, 4}, {1, 2, 3, 6, 7}}); set_cohs(ce, {{4, 5, 6, 7}}); }
            else { set_subs(ce, {{0, 7, 6}, {1, 2, 3, 4, 5}}); set_cohs(ce, {{5, 4, 7, 6}}); }
        } else if (n0e2 && n0e3) {   // node 1 separate
            set_edges(ce, {0, 1}); sm[0] = M(n1); sm[1] = M(n3);
            if (!coh) set_subs(ce, {{1, 5, 4}, {2, 3, 0, 4, 5}});
            else if (n1 < n0) { set_subs(ce, {{1, 5, 4}, {2, 3, 0, 6, 7}}); set_cohs(ce, {{4, 5, 6, 7}}); }
            else { set_subs(ce, {{1, 7, 6}, {2, 3, 0, 4, 5}}); set_cohs(ce, {{5, 4, 7, 6}}); }
        } else if (n0e1 && n0e3) {   // node 2 separate
            set_edges(ce, {1, 2}); sm[0] = M(n2); sm[1] = M(n0);
            if (!coh) set_subs(ce, {{2, 5, 4}, {3, 0, 1, 4, 5}});
            else if (n2 < n0) { set_subs(ce, {{2, 5, 4}, {3, 0, 1, 6, 7}}); set_cohs(ce, {{4, 5, 6, 7}}); }
            else { set_subs(ce, {{2, 7, 6}, {3, 0, 1, 4, 5}}); set_cohs(ce, {{5, 4, 7, 6}}); }
        } else if (n0e1 && n1e2) {   // node 3 separate
            set_edges(ce, {2, 3}); sm[0] = M(n3); sm[1] = M(n1);
            if (!coh) set_subs(ce, {{3, 5, 4}, {0, 1, 2, 4, 5}});
            else if (n3 < n0) { set_subs(ce, {{3, 5, 4}, {0, 1, 2, 6, 7}}); set_cohs(ce, {{4, 5, 6, 7}}); }
            else { set_subs(ce, {{3, 7, 6}, {0, 1, 2, 4, 5}}); set_cohs(ce, {{5, 4, 7, 6}}); }
        } else fail(PGF_ERR_UNSUPPORTED, "quadrilateral has 1 node separated in an unknown manner");
    } else {  // two and two
        if (n0e1) {
            set_edges(ce, {1, 3});
            if (!coh) { sm[0] = M(n0); sm[1] = M(n2); set_subs(ce, {{0, 1, 4, 5}, {2, 3, 5, 4}}); }
            else {
                sm[0] = M(n0); sm[1] = M(n2);
                if (n0 < n2) { set_subs(ce, {{0, 1, 4, 5}, {2, 3, 7, 6}}); set_cohs(ce, {{5, 4, 7, 6}}); }
                else { set_subs(ce, {{0, 1, 6, 7}, {2, 3, 5, 4}}); set_cohs(ce, {{4, 5, 6, 7}}); }
            }
        } else if (n1e2) {
            set_edges(ce, {2, 0});
            if (!coh) { sm[0] = M(n1); sm[1] = M(n3); set_subs(ce, {{1, 2, 4, 5}, {3, 0, 5, 4}}); }
            else {
                sm[0] = M(n0); sm[1] = M(n1);
                if (n0 < n1) { set_subs(ce, {{3, 0, 5, 4}, {1, 2, 6, 7}}); set_cohs(ce, {{4, 5, 6, 7}}); }
                else { set_subs(ce, {{3, 0, 7, 6}, {1, 2, 4, 5}}); set_cohs(ce, {{5, 4, 7, 6}}); }
            }
        } else {
            std::ostringstream s;
            s << "quadrilateral element " << e << " has 2 nodes separated in an unknown manner (refine)";
            fail(PGF_ERR_UNSUPPORTED, s.str());
        }
    }
    for (int i = 0; i < ce.n_sub; ++i) ce.sub_mat[i] = sm[i];
    for (int i = 0; i < ce.n_cut; ++i) ce.enrich_incl[i] = vec[nu - 1];
    ce.n_enrich = (uint8_t)(coh ? 2 * ce.n_cut : ce.n_cut);
}

// src/tet4.cpp:228-520
void Mesh::detect_tet4(int64_t e, const int* nd, CutElem& ce, bool& cut) {
    const int n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
    const bool n0e1 = n0 == n1, n0e2 = n0 == n2, n0e3 = n0 == n3, n1e2 = n1 == n2, n1e3 = n1 == n3,
               n2e3 = n2 == n3;
    int vec[4], cnt[4], nu;
    unique_counts(nd, 4, vec, nu, cnt);
    cut = nu > 1;
    if (!cut) return;
    if (nu != 2 || vec[0] >= 0) {
        std::ostringstream s;
        s << "element " << e << " is cut by more than one interface (unsupported)";
        fail(PGF_ERR_UNSUPPORTED, s.str());
    }
    auto M = [&](int v) { return v < 0 ? _elem_mat[e] : _incl[v].mat; };
    const bool coh = _opt.cohesive;
    int a = 0, b = 0;  // materials of the first and second group of sub-elements
    int ngroup_a = 1;
    if (cnt[0] == 1 || cnt[0] == 3) {
        if (n1e2 && n1e3) {
            set_edges(ce, {0, 3, 2}); a = M(n0); b = M(n1);
            if (!coh) set_subs(ce, {{4, 5, 6, 0}, {1, 3, 2, 4}, {2, 4, 3, 5}, {4, 6, 5, 2}});
            else if (n0 < n1) { set_subs(ce, {{4, 5, 6, 0}, {1, 3, 2, 7}, {2, 7, 3, 8}, {7, 9, 8, 2}}); set_cohs(ce, {{4, 6, 5, 7, 9, 8}}); }
            else { set_subs(ce, {{7, 8, 9, 0}, {1, 3, 2, 4}, {2, 4, 3, 5}, {4, 6, 5, 2}}); set_cohs(ce, {{4, 5, 6, 7, 8, 9}}); }
        } else if (n0e2 && n0e3) {
            set_edges(ce, {0, 1, 4}); a = M(n1); b = M(n0);
            if (!coh) set_subs(ce, {{4, 5, 6, 1}, {0, 2, 3, 4}, {2, 3, 4, 5}, {4, 6, 5, 3}});
            else if (n1 < n0) { set_subs(ce, {{4, 5, 6, 1}, {0, 2, 3, 7}, {2, 3, 7, 8}, {7, 9, 8, 3}}); set_cohs(ce, {{4, 6, 5, 7, 9, 8}}); }
            else { set_subs(ce, {{7, 8, 9, 1}, {0, 2, 3, 4}, {2, 3, 4, 5}, {4, 6, 5, 3}}); set_cohs(ce, {{4, 5, 6, 7, 8, 9}}); }
        } else if (n0e1 && n0e3) {
            set_edges(ce, {1, 2, 5}); a = M(n2); b = M(n1);
            if (!coh) set_subs(ce, {{4, 5, 6, 2}, {1, 0, 3, 4}, {0, 3, 4, 5}, {4, 6, 5, 3}});
            else if (n2 < n0) { set_subs(ce, {{4, 5, 6, 2}, {0, 1, 7, 3}, {0, 3, 7, 8}, {7, 9, 8, 3}}); set_cohs(ce, {{4, 6, 5, 7, 9, 8}}); }
            else { set_subs(ce, {{7, 8, 9, 2}, {0, 1, 4, 3}, {0, 3, 4, 5}, {4, 6, 5, 3}}); set_cohs(ce, {{4, 5, 6, 7, 8, 9}}); }
        } else if (n0e1 && n1e2) {
            set_edges(ce, {3, 4, 5}); a = M(n3); b = M(n1);
            if (!coh) set_subs(ce, {{4, 5, 6, 3}, {0, 1, 2, 4}, {1, 2, 4, 5}, {4, 6, 5, 2}});
            else if (n3 < n1) { set_subs(ce, {{4, 5, 6, 3}, {0, 1, 2, 7}, {1, 2, 7, 8}, {7, 9, 8, 2}}); set_cohs(ce, {{4, 6, 5, 7, 9, 8}}); }
            else { set_subs(ce, {{7, 8, 9, 3}, {0, 1, 2, 4}, {1, 2, 4, 5}, {4, 6, 5, 2}}); set_cohs(ce, {{4, 5, 6, 7, 8, 9}}); }
        } else fail(PGF_ERR_UNSUPPORTED, "tetrahedron has 1 node separated in an unknown manner");
        ngroup_a = 1;
    } else {
        if (n0e1 && n2e3 && !n0e2) {
            set_edges(ce, {1, 2, 3, 4}); a = M(n0); b = M(n2);
            if (!coh) set_subs(ce, {{0, 1, 4, 7}, {0, 7, 4, 6}, {0, 6, 4, 5}, {2, 6, 7, 3}, {2, 6, 4, 7}, {2, 5, 4, 6}});
            else if (n0 < n2) { set_subs(ce, {{0, 1, 4, 7}, {0, 7, 4, 6}, {0, 6, 4, 5}, {2, 10, 11, 3}, {2, 10, 8, 11}, {2, 9, 8, 10}}); set_cohs(ce, {{4, 5, 6, 7, 8, 9, 10, 11}}); }
            else { set_subs(ce, {{0, 1, 8, 11}, {0, 11, 8, 10}, {0, 10, 8, 9}, {2, 6, 7, 3}, {2, 6, 4, 7}, {2, 5, 4, 6}}); set_cohs(ce, {{7, 6, 5, 4, 11, 10, 9, 8}}); }
        } else if (n1e2 && n0e3 && !n0e1) {
            set_edges(ce, {0, 4, 5, 2}); a = M(n1); b = M(n0);
            if (!coh) set_subs(ce, {{1, 2, 4, 5}, {2, 7, 4, 5}, {2, 7, 5, 6}, {0, 4, 7, 3}, {4, 5, 7, 3}, {3, 5, 7, 6}});
            else if (n1 < n3) { set_subs(ce, {{1, 2, 4, 5}, {2, 7, 4, 5}, {2, 7, 5, 6}, {0, 8, 11, 3}, {8, 9, 11, 3}, {3, 9, 11, 10}}); set_cohs(ce, {{7, 6, 5, 4, 11, 10, 9, 8}}); }
            else { set_subs(ce, {{1, 2, 8, 9}, {2, 11, 8, 9}, {2, 11, 9, 10}, {0, 4, 7, 3}, {4, 5, 7, 3}, {3, 5, 7, 6}}); set_cohs(ce, {{4, 5, 6, 7, 8, 9, 10, 11}}); }
        } else if (n0e2 && n1e3 && !n0e1) {
            set_edges(ce, {0, 1, 5, 3}); a = M(n0); b = M(n1);
            if (!coh) set_subs(ce, {{0, 4, 2, 7}, {2, 7, 4, 6}, {2, 4, 5, 6}, {1, 6, 7, 3}, {1, 6, 4, 7}, {1, 5, 4, 6}});
            else if (n0 < n1) { set_subs(ce, {{0, 4, 2, 7}, {2, 7, 4, 6}, {2, 4, 5, 6}, {1, 10, 11, 3}, {1, 10, 8, 11}, {1, 9, 8, 10}}); set_cohs(ce, {{4, 5, 6, 7, 8, 9, 10, 11}}); }
            else { set_subs(ce, {{0, 8, 2, 11}, {2, 11, 8, 10}, {2, 8, 9, 10}, {1, 6, 7, 3}, {1, 6, 4, 7}, {1, 5, 4, 6}}); set_cohs(ce, {{7, 6, 5, 4, 11, 10, 9, 8}}); }
        } else {
            std::ostringstream s;
            s << "tetrahedron " << e << " has 2 nodes separated in an unknown manner (refine)";
            fail(PGF_ERR_UNSUPPORTED, s.str());
        }
        ngroup_a = 3;
    }
    for (int i = 0; i < ce.n_sub; ++i) ce.sub_mat[i] = i < ngroup_a ? a : b;
    for (int i = 0; i < ce.n_cut; ++i) ce.enrich_incl[i] = vec[nu - 1];
    ce.n_enrich = (uint8_t)(coh ? 2 * ce.n_cut : ce.n_cut);
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
        if (_kind == QUAD4) detect_quad4(e, nd, ce, cut);
        else detect_tet4(e, nd, ce, cut);
        if (!cut) {
            if (nd[0] != -1) _elem_mat[e] = _incl[nd[0]].mat;
            continue;
        }
        ce.elem = (int32_t)e;
        for (int c = 0; c < ce.n_coh; ++c) {
            // every cohesive node is an enrichment node; its inclusion picks the law
            int slot = ce.coh_nodes[c][0] - 4;
            int inc = ce.enrich_incl[slot % ce.n_cut];
            ce.coh_law[c] = _incl[inc].coh_law >= 0 ? _incl[inc].coh_law : _opt.default_coh_law;
        }
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
        EdgeLevelSet ls;
        ls.inc = &_incl[keys[k][2]];
        ls.nd = ls.inc->type == 0 ? 2 : 3;
        for (int d = 0; d < 3; ++d) { ls.p0[d] = _X[d][keys[k][0]]; ls.p1[d] = _X[d][keys[k][1]]; }
        double a = find_root(ls);
        for (int d = 0; d < 3; ++d) {
            double v = ls.p0[d] + a * (ls.p1[d] - ls.p0[d]);
            _X[d][_n_nodes + k] = v;
            if (_opt.cohesive) _X[d][_n_nodes + n_init + k] = v;
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
namespace {

void parent_inverse_map(ElemKind kind, const double (*P)[3], const double* g, double* xi) {
    if (kind == TET4) {
        double J[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) J[i][j] = P[j + 1][i] - P[0][i];
        double r[3] = {g[0] - P[0][0], g[1] - P[0][1], g[2] - P[0][2]};
        double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        // Cramer's rule
        for (int c = 0; c < 3; ++c) {
            double M[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) M[i][j] = j == c ? r[i] : J[i][j];
            double dc = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                        M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
            xi[c] = dc / det;
        }
        return;
    }
    // bilinear quad: Newton from the origin (exact after one step for parallelograms)
    xi[0] = 0.0; xi[1] = 0.0;
    for (int it = 0; it < 50; ++it) {
        double N[4], dN[4][3];
        geom::quad4_shape(xi, N, dN);
        double f[2] = {-g[0], -g[1]}, J[2][2] = {{0, 0}, {0, 0}};
        for (int n = 0; n < 4; ++n)
            for (int i = 0; i < 2; ++i) {
                f[i] += N[n] * P[n][i];
                for (int j = 0; j < 2; ++j) J[i][j] += P[n][i] * dN[n][j];
            }
        double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        double d0 = (J[1][1] * f[0] - J[0][1] * f[1]) / det, d1 = (-J[1][0] * f[0] + J[0][0] * f[1]) / det;
        xi[0] -= d0; xi[1] -= d1;
        if (std::fabs(d0) + std::fabs(d1) < 1e-15) break;
    }
}

bool parent_inside(ElemKind kind, const double* xi) {
    const double tol = 1e-10;  // Elem::_inside_tol (src/elem.cpp:23)
    if (kind == QUAD4) return xi[0] >= -1.0 - tol && xi[0] <= 1.0 + tol && xi[1] >= -1.0 - tol && xi[1] <= 1.0 + tol;
    return xi[0] >= -tol && xi[0] <= 1.0 + tol && xi[1] >= -tol && xi[1] <= 1.0 + tol && xi[2] >= -tol &&
           xi[2] <= 1.0 + tol && xi[0] + xi[1] + xi[2] <= 1.0 + tol;
}

}  // namespace

void Mesh::store_shape_functions() {
    if (!_enriched && !_cut.empty()) fail(PGF_ERR_STATE, "add_enrichments must run before store_shape_functions");
    const int dim = _dim;
    vpt_ptr.assign(1, 0); vpt_w.clear(); vpt_mat.clear(); vpt_grad_off.clear(); vpt_grad.clear();
    cpt_ptr.assign(1, 0); cpt_w.clear(); cpt_N.clear(); cpt_R.clear(); cpt_nodes.clear(); cpt_nh.clear(); cpt_law.clear();
    for (auto& ce : _cut) {
        const int64_t e = ce.elem;
        const int nt = 4 + ce.n_enrich;
        double X[MAX_EL_NODES][3];
        for (int s = 0; s < nt; ++s) {
            int32_t g = elem_node(e, s);
            for (int d = 0; d < 3; ++d) X[s][d] = _X[d][g];
        }
        // parent gradients are needed at every child point
        for (int ie = 0; ie < ce.n_sub; ++ie) {
            const int cn = ce.sub_nn[ie];
            double CX[8][3];
            for (int k = 0; k < cn; ++k)
                for (int d = 0; d < 3; ++d) CX[k][d] = X[ce.sub_nodes[ie][k]][d];
            int nqp;
            if (dim == 3) nqp = geom::kTet4NQP;
            else if (cn == 3) nqp = geom::kTri3NQP;
            else if (cn == 4) nqp = geom::kQuad4NQP;
            else nqp = cn * geom::kTri3NQP;
            double wsum = 0.0;
            double grad[MAX_EL_NODES][3];
            for (int qp = 0; qp < nqp; ++qp) {
                double xi[3] = {0, 0, 0}, W;
                double cN[8], cdN[8][3], cdx[8][3];
                if (dim == 3) { geom::tet4_qp(qp, xi, W); geom::tet4_shape(xi, cN, cdN); }
                else if (cn == 3) { geom::tri3_qp(qp, xi, W); geom::tri3_shape(xi, cN, cdN); }
                else if (cn == 4) { geom::quad4_qp(qp, xi, W); geom::quad4_shape(xi, cN, cdN); }
                else { geom::polygon_qp(cn, qp, xi, W); geom::polygon_shape(cn, xi, cN, cdN); }
                double Jc = geom::transform_gradients(dim, cn, CX, cdN, cdx);
                if (Jc < 0.0) {
                    std::ostringstream s;
                    s << "negative sub-element Jacobian in element " << e;
                    fail(PGF_ERR_GEOMETRY, s.str());
                }
                double gpt[3] = {0, 0, 0};
                for (int d = 0; d < dim; ++d)
                    for (int k = 0; k < cn; ++k) gpt[d] += CX[k][d] * cN[k];
                double pxi[3] = {0, 0, 0};
                parent_inverse_map(_kind, X, gpt, pxi);
                if (!parent_inside(_kind, pxi)) fail(PGF_ERR_GEOMETRY, "bad inverse map detected");
                double pN[4], pdN[4][3], pdx[4][3];
                if (_kind == QUAD4) geom::quad4_shape(pxi, pN, pdN);
                else geom::tet4_shape(pxi, pN, pdN);
                geom::transform_gradients(dim, 4, X, pdN, pdx);
                double chk = 0.0;
                for (int d = 0; d < dim; ++d) {
                    double s = 0.0;
                    for (int n = 0; n < 4; ++n) s += X[n][d] * pN[n];
                    chk += (s - gpt[d]) * (s - gpt[d]);
                }
                if (std::sqrt(chk) > 1e-9) fail(PGF_ERR_GEOMETRY, "bad inverse map in assembly detected");
                for (int s = 0; s < nt; ++s)
                    for (int d = 0; d < 3; ++d) grad[s][d] = 0.0;
                for (int n = 0; n < 4; ++n)
                    for (int d = 0; d < dim; ++d) grad[n][d] = pdx[n][d];
                for (int en = 0; en < ce.n_enrich; ++en)
                    for (int k = 0; k < cn; ++k)
                        if (ce.sub_nodes[ie][k] == 4 + en) {
                            for (int d = 0; d < dim; ++d) grad[4 + en][d] = cdx[k][d];
                            break;
                        }
                double w = Jc * W;
                if (dim == 3) { wsum += w; continue; }
                vpt_w.push_back(w);
                vpt_mat.push_back(ce.sub_mat[ie]);
                vpt_grad_off.push_back((int64_t)vpt_grad.size());
                for (int s = 0; s < nt; ++s)
                    for (int d = 0; d < dim; ++d) vpt_grad.push_back(grad[s][d]);
            }
            if (dim == 3) {
                vpt_w.push_back(wsum * _opt.tet_weight_scale);
                vpt_mat.push_back(ce.sub_mat[ie]);
                vpt_grad_off.push_back((int64_t)vpt_grad.size());
                for (int s = 0; s < nt; ++s)
                    for (int d = 0; d < dim; ++d) vpt_grad.push_back(grad[s][d]);
            }
        }
        vpt_ptr.push_back((int64_t)vpt_w.size());
        for (int c = 0; c < ce.n_coh; ++c) {
            const int nh = ce.coh_nn[c] / 2;
            const int sd = dim - 1;  // surface dimension
            int nqp = dim == 2 ? geom::kEdge2NQP : (nh == 3 ? geom::kTri3NQP : geom::kQuad4NQP);
            for (int qp = 0; qp < nqp; ++qp) {
                double xi[2] = {0, 0}, W, N[4], dN[4][3];
                if (dim == 2) {
                    geom::edge2_qp(qp, xi[0], W);
                    N[0] = 0.5 * (1.0 - xi[0]); N[1] = 0.5 * (1 + xi[0]);
                    dN[0][0] = -0.5; dN[1][0] = 0.5;
                } else if (nh == 3) {
                    geom::tri3_qp(qp, xi, W);
                    geom::tri3_shape(xi, N, dN);
                } else {
                    geom::quad4_qp(qp, xi, W);
                    geom::quad4_shape(xi, N, dN);
                }
                double Jt[2][3] = {{0, 0, 0}, {0, 0, 0}};
                for (int n = 0; n < nh; ++n) {
                    const double* xn = X[ce.coh_nodes[c][n]];
                    for (int a = 0; a < sd; ++a)
                        for (int i = 0; i < dim; ++i) Jt[a][i] += dN[n][a] * xn[i];
                }
                double nrm[3];
                if (dim == 2) { nrm[0] = -Jt[0][1]; nrm[1] = Jt[0][0]; nrm[2] = 0; }
                else {
                    nrm[0] = Jt[0][1] * Jt[1][2] - Jt[0][2] * Jt[1][1];
                    nrm[1] = Jt[0][2] * Jt[1][0] - Jt[0][0] * Jt[1][2];
                    nrm[2] = Jt[0][0] * Jt[1][1] - Jt[0][1] * Jt[1][0];
                }
                double dA = 0.0;
                for (int i = 0; i < dim; ++i) dA += nrm[i] * nrm[i];
                dA = std::sqrt(dA);
                double R[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int i = 0; i < dim; ++i) R[(dim - 1) * dim + i] = nrm[i] / dA;
                for (int a = 0; a < sd; ++a) {
                    double mag = 0.0;
                    for (int i = 0; i < dim; ++i) mag += Jt[a][i] * Jt[a][i];
                    mag = std::sqrt(mag);
                    for (int i = 0; i < dim; ++i) R[a * dim + i] = Jt[a][i] / mag;
                }
                cpt_w.push_back(dA * W);
                for (int n = 0; n < 4; ++n) cpt_N.push_back(n < nh ? N[n] : 0.0);
                for (int k = 0; k < 9; ++k) cpt_R.push_back(R[k]);
                for (int k = 0; k < 8; ++k) cpt_nodes.push_back(k < 2 * nh ? ce.coh_nodes[c][k] : 0);
                cpt_nh.push_back((uint8_t)nh);
                cpt_law.push_back(ce.coh_law[c]);
            }
        }
        cpt_ptr.push_back((int64_t)cpt_w.size());
    }
}

}  // namespace pgf
