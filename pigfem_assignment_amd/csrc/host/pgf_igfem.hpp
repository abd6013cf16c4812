// IGFEM preprocessing steps that run both on the host (Mesh::analyze_nodes,
// analyze_elements, add_enrichments, store_shape_functions) and in the device
// set-up kernels: one source, so the two paths give the same flags, the same
// enrichment coordinates and the same integration tables bit for bit (the
// device code is built without multiply-add contraction; +, -, *, / and sqrt
// round correctly on both sides).
//
// Reference routines: Mesh::detect_node and the inclusion level sets
// (src/Mesh.cpp:2838-2891, src/InclusionCircle.cpp:150-163,
// src/InclusionSphere.cpp:163-178), the element detection tables
// (src/quad4.cpp:188-560, src/tet4.cpp:228-520), Inclusion::find_intersection
// (src/Inclusion.cpp:32-123) with GSL 2.1's Steffenson and Brent steps, and
// Mesh::store_shape_functions (src/Mesh.cpp:1443-1605,
// src/CohesiveElem.cpp:130-242).
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <initializer_list>

#include "pgf_common.hpp"
#include "pgf_geom.hpp"
#include "pgf_mesh.hpp"

namespace pgf {
namespace igfem {

// ---------------------------------------------------------------------------
// Node detection.  Scans the inclusions last to first (the last one added
// wins).  Returns j >= 0 when the point is inside inclusion j, -1 when it is
// in the matrix, and -2 - j when it lies within 1e-12 (relative) of inclusion
// j's surface, where the caller moves the node and scans again.
PGF_HD inline int scan_node(const Inclusion* incl, int ni, bool zlim_as_reference, double x, double y, double z) {
    for (int j = ni - 1; j >= 0; --j) {
        const Inclusion& inc = incl[j];
        double lo[3], hi[3];
        if (inc.type == 0) {
            lo[0] = inc.c[0] - inc.r; lo[1] = inc.c[1] - inc.r; lo[2] = -DBL_MAX;
            hi[0] = inc.c[0] + inc.r; hi[1] = inc.c[1] + inc.r; hi[2] = DBL_MAX;
        } else {
            lo[0] = inc.c[0] - inc.r; lo[1] = inc.c[1] - inc.r; lo[2] = inc.c[2] - inc.r;
            hi[0] = inc.c[0] + inc.r; hi[1] = inc.c[1] + inc.r;
            hi[2] = zlim_as_reference ? inc.c[1] + inc.r : inc.c[2] + inc.r;
        }
        if (!(lo[0] <= x && lo[1] <= y && lo[2] <= z && x <= hi[0] && y <= hi[1] && z <= hi[2])) continue;
        double val;
        if (inc.type == 0) {
            double gx = x - inc.c[0], gy = y - inc.c[1];
            val = (sqrt(gx * gx + gy * gy) - inc.r) / inc.r;
        } else {
            double gx = (x - inc.c[0]) / inc.r, gy = (y - inc.c[1]) / inc.r, gz = (z - inc.c[2]) / inc.r;
            val = gx * gx + gy * gy + gz * gz - 1.0;
        }
        const double eps = 1e-12;
        if (val > eps) continue;
        if (val < -eps) return j;
        return -2 - j;
    }
    return -1;
}

// ---------------------------------------------------------------------------
// Element detection tables.
enum ClassifyStatus : int {
    CLS_OK = 0,
    CLS_MULTI = 1,        // touched by more than one inclusion
    CLS_ONE_UNKNOWN = 2,  // one node separated in a pattern the tables do not hold
    CLS_TWO_UNKNOWN = 3,  // two and two separated in a pattern the tables do not hold
};

// element edge -> local nodes (src/quad4.cpp:23-29, src/tet4.cpp:22-30)
PGF_HD inline void edge_nodes(int kind, int edge, int& a, int& b) {
    constexpr int kQuadEdge[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    constexpr int kTetEdge[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    if (kind == QUAD4) { a = kQuadEdge[edge][0]; b = kQuadEdge[edge][1]; }
    else { a = kTetEdge[edge][0]; b = kTetEdge[edge][1]; }
}

PGF_HD inline void set_subs(CutElem& ce, std::initializer_list<std::initializer_list<int>> subs) {
    ce.n_sub = 0;
    for (auto& s : subs) {
        int k = 0;
        for (int v : s) ce.sub_nodes[ce.n_sub][k++] = (uint8_t)v;
        ce.sub_nn[ce.n_sub++] = (uint8_t)k;
    }
}
PGF_HD inline void set_cohs(CutElem& ce, std::initializer_list<std::initializer_list<int>> cohs) {
    ce.n_coh = 0;
    for (auto& s : cohs) {
        int k = 0;
        for (int v : s) ce.coh_nodes[ce.n_coh][k++] = (uint8_t)v;
        ce.coh_nn[ce.n_coh++] = (uint8_t)k;
    }
}
PGF_HD inline void set_edges(CutElem& ce, std::initializer_list<int> edges) {
    ce.n_cut = 0;
    for (int v : edges) ce.cut_edge[ce.n_cut++] = (uint8_t)v;
}

// distinct values of nd[0..3] in increasing order and how often each occurs
PGF_HD inline void unique_counts(const int* nd, int* vec, int& nu, int* counts) {
    int t[4] = {nd[0], nd[1], nd[2], nd[3]};
    for (int i = 1; i < 4; ++i)
        for (int j = i; j > 0 && t[j] < t[j - 1]; --j) { int s = t[j]; t[j] = t[j - 1]; t[j - 1] = s; }
    nu = 0;
    for (int i = 0; i < 4; ++i)
        if (nu == 0 || vec[nu - 1] != t[i]) vec[nu++] = t[i];
    for (int i = 0; i < nu; ++i) {
        counts[i] = 0;
        for (int k = 0; k < 4; ++k) counts[i] += nd[k] == vec[i];
    }
}

// src/quad4.cpp:188-560.  nd: detection values of the element's nodes,
// emat: the element's material before detection.
PGF_HD inline int classify_quad4(const int* nd, int emat, const Inclusion* incl, bool coh, CutElem& ce, bool& cut) {
    const int n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
    const bool n0e1 = n0 == n1, n0e2 = n0 == n2, n0e3 = n0 == n3, n1e2 = n1 == n2, n1e3 = n1 == n3;
    int vec[4], cnt[4], nu;
    unique_counts(nd, vec, nu, cnt);
    cut = nu > 1;
    if (!cut) return CLS_OK;
    if (nu != 2 || vec[0] >= 0) return CLS_MULTI;
    auto M = [&](int v) { return v < 0 ? emat : incl[v].mat; };
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
        } else return CLS_ONE_UNKNOWN;
    } else {  // two and two
        if (n0e1) {
            set_edges(ce, {1, 3});
            sm[0] = M(n0); sm[1] = M(n2);
            if (!coh) set_subs(ce, {{0, 1, 4, 5}, {2, 3, 5, 4}});
            else if (n0 < n2) { set_subs(ce, {{0, 1, 4, 5}, {2, 3, 7, 6}}); set_cohs(ce, {{5, 4, 7, 6}}); }
            else { set_subs(ce, {{0, 1, 6, 7}, {2, 3, 5, 4}}); set_cohs(ce, {{4, 5, 6, 7}}); }
        } else if (n1e2) {
            set_edges(ce, {2, 0});
            if (!coh) { sm[0] = M(n1); sm[1] = M(n3); set_subs(ce, {{1, 2, 4, 5}, {3, 0, 5, 4}}); }
            else {
                sm[0] = M(n0); sm[1] = M(n1);
                if (n0 < n1) { set_subs(ce, {{3, 0, 5, 4}, {1, 2, 6, 7}}); set_cohs(ce, {{4, 5, 6, 7}}); }
                else { set_subs(ce, {{3, 0, 7, 6}, {1, 2, 4, 5}}); set_cohs(ce, {{5, 4, 7, 6}}); }
            }
        } else return CLS_TWO_UNKNOWN;
    }
    for (int i = 0; i < ce.n_sub; ++i) ce.sub_mat[i] = sm[i];
    for (int i = 0; i < ce.n_cut; ++i) ce.enrich_incl[i] = vec[nu - 1];
    ce.n_enrich = (uint8_t)(coh ? 2 * ce.n_cut : ce.n_cut);
    return CLS_OK;
}

// src/tet4.cpp:228-520
PGF_HD inline int classify_tet4(const int* nd, int emat, const Inclusion* incl, bool coh, CutElem& ce, bool& cut) {
    const int n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
    const bool n0e1 = n0 == n1, n0e2 = n0 == n2, n0e3 = n0 == n3, n1e2 = n1 == n2, n1e3 = n1 == n3,
               n2e3 = n2 == n3;
    int vec[4], cnt[4], nu;
    unique_counts(nd, vec, nu, cnt);
    cut = nu > 1;
    if (!cut) return CLS_OK;
    if (nu != 2 || vec[0] >= 0) return CLS_MULTI;
    auto M = [&](int v) { return v < 0 ? emat : incl[v].mat; };
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
        } else return CLS_ONE_UNKNOWN;
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
        } else return CLS_TWO_UNKNOWN;
        ngroup_a = 3;
    }
    for (int i = 0; i < ce.n_sub; ++i) ce.sub_mat[i] = i < ngroup_a ? a : b;
    for (int i = 0; i < ce.n_cut; ++i) ce.enrich_incl[i] = vec[nu - 1];
    ce.n_enrich = (uint8_t)(coh ? 2 * ce.n_cut : ce.n_cut);
    return CLS_OK;
}

// every cohesive node is an enrichment node; its inclusion picks the law
PGF_HD inline void cohesive_laws(CutElem& ce, const Inclusion* incl, int default_law) {
    for (int c = 0; c < ce.n_coh; ++c) {
        int slot = ce.coh_nodes[c][0] - 4;
        int inc = ce.enrich_incl[slot % ce.n_cut];
        ce.coh_law[c] = incl[inc].coh_law >= 0 ? incl[inc].coh_law : default_law;
    }
}

// ---------------------------------------------------------------------------
// Enrichment-node coordinates: root of the inclusion's level set along the
// edge p0 + a (p1 - p0), as the reference's int_f / int_df
// (src/InclusionCircle.cpp:248-287, src/InclusionSphere.cpp:270-326).
struct EdgeLevelSet {
    double c[3], r;
    double p0[3], p1[3];
    int nd;  // 2 for circles, 3 for spheres
    PGF_HD double f(double a) const {
        double s = 0.0;
        for (int d = 0; d < nd; ++d) {
            double g = (p0[d] + a * (p1[d] - p0[d]) - c[d]) / r;
            s += g * g;
        }
        return s - 1.0;
    }
    PGF_HD double df(double a) const {
        double s = 0.0;
        for (int d = 0; d < nd; ++d) {
            double g = p0[d] + a * (p1[d] - p0[d]) - c[d];
            s += g * (p1[d] - p0[d]);
        }
        return (1.0 / (r * r)) * 2.0 * s;
    }
};

PGF_HD inline double dmin(double a, double b) { return b < a ? b : a; }
PGF_HD inline double dmax(double a, double b) { return a < b ? b : a; }

// GSL 2.1 Steffenson iteration (roots/steffenson.c): Newton steps, with the
// Aitken delta-squared estimate reported from the third iterate on.
// GSL 2.1 Brent (roots/brent.c) as the bracketing fallback.  Driven exactly as
// Inclusion::find_intersection drives them (src/Inclusion.cpp:32-123):
// start 0.5, |delta| < 1e-7, 25 iterations, Brent on [0,1] when Steffenson
// runs out of iterations or leaves [0,1].  Returns false when the edge holds
// no sign change.
PGF_HD inline bool find_root(const EdgeLevelSet& ls, double& out) {
    if (ls.f(0.0) * ls.f(1.0) >= 0) return false;
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
        conv = fabs(alpha - alpha0) < 1e-7 || alpha == alpha0;
    } while (!conv && iter < max_iter);
    if (iter == max_iter || alpha < 0.0 || alpha > 1.0) {
        // Brent
        double a = 0.0, b = 1.0, c = 1.0, fa = ls.f(0.0), fb = ls.f(1.0), fc = fb;
        double d = 1.0, e = 1.0, lo = 0.0, hi = 1.0;
        alpha = 0.5;
        for (int it = 0; it < 100; ++it) {
            bool ac_equal = false;
            if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) { ac_equal = true; c = a; fc = fa; d = b - a; e = b - a; }
            if (fabs(fc) < fabs(fb)) { ac_equal = true; a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
            double tol = 0.5 * DBL_EPSILON * fabs(b), m = 0.5 * (c - b);
            if (fb == 0) { alpha = b; lo = hi = b; }
            else if (fabs(m) <= tol) { alpha = b; lo = dmin(b, c); hi = dmax(b, c); }
            else {
                if (fabs(e) < tol || fabs(fa) <= fabs(fb)) { d = m; e = m; }
                else {
                    double p, q, r, s = fb / fa;
                    if (ac_equal) { p = 2 * m * s; q = 1 - s; }
                    else {
                        q = fa / fc; r = fb / fc;
                        p = s * (2 * m * q * (q - r) - (b - a) * (r - 1));
                        q = (q - 1) * (r - 1) * (s - 1);
                    }
                    if (p > 0) q = -q; else p = -p;
                    if (2 * p < dmin(3 * m * q - fabs(tol * q), fabs(e * q))) { e = d; d = p / q; }
                    else { d = m; e = m; }
                }
                a = b; fa = fb;
                if (fabs(d) > tol) b += d; else b += (m > 0 ? +tol : -tol);
                fb = ls.f(b);
                alpha = b;
                double cc = c;
                if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) cc = a;
                lo = dmin(b, cc); hi = dmax(b, cc);
            }
            if (fabs(lo - hi) < 1e-7 || lo == hi) break;
        }
    }
    out = alpha;
    return true;
}

// Enrichment node of the edge (n0, n1) of inclusion inc: the level-set root
// interpolated between the endpoint coordinates.
PGF_HD inline bool enrichment_point(const Inclusion& inc, const double* p0, const double* p1, double* x) {
    EdgeLevelSet ls;
    ls.nd = inc.type == 0 ? 2 : 3;
    for (int d = 0; d < 3; ++d) { ls.c[d] = inc.c[d]; ls.p0[d] = p0[d]; ls.p1[d] = p1[d]; }
    ls.r = inc.r;
    double a;
    if (!find_root(ls, a)) return false;
    for (int d = 0; d < 3; ++d) x[d] = ls.p0[d] + a * (ls.p1[d] - ls.p0[d]);
    return true;
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

PGF_HD inline void parent_inverse_map(int kind, const double (*P)[3], const double* g, double* xi) {
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
        if (fabs(d0) + fabs(d1) < 1e-15) break;
    }
}

PGF_HD inline bool parent_inside(int kind, const double* xi) {
    const double tol = 1e-10;  // Elem::_inside_tol (src/elem.cpp:23)
    if (kind == QUAD4) return xi[0] >= -1.0 - tol && xi[0] <= 1.0 + tol && xi[1] >= -1.0 - tol && xi[1] <= 1.0 + tol;
    return xi[0] >= -tol && xi[0] <= 1.0 + tol && xi[1] >= -tol && xi[1] <= 1.0 + tol && xi[2] >= -tol &&
           xi[2] <= 1.0 + tol && xi[0] + xi[1] + xi[2] <= 1.0 + tol;
}

PGF_HD inline int sub_nqp(int dim, int cn) {
    if (dim == 3) return geom::kTet4NQP;
    if (cn == 3) return geom::kTri3NQP;
    if (cn == 4) return geom::kQuad4NQP;
    return cn * geom::kTri3NQP;
}

// Number of volumetric and cohesive table points of a cut element.
PGF_HD inline void table_sizes(int dim, const CutElem& ce, int& nv, int& nc) {
    nv = 0;
    for (int ie = 0; ie < ce.n_sub; ++ie) nv += dim == 3 ? 1 : sub_nqp(dim, ce.sub_nn[ie]);
    nc = 0;
    for (int c = 0; c < ce.n_coh; ++c) {
        const int nh = ce.coh_nn[c] / 2;
        nc += dim == 2 ? geom::kEdge2NQP : (nh == 3 ? geom::kTri3NQP : geom::kQuad4NQP);
    }
}

enum TableStatus : int { TAB_OK = 0, TAB_NEG_JACOBIAN = 1, TAB_OUTSIDE = 2, TAB_MAP_CHECK = 3 };

// Where one cut element's table entries go (each pointer at the element's
// first entry; vgoff values are global offsets starting at goff0).
struct TableOut {
    double* vw;
    int32_t* vmat;
    int64_t* vgoff;
    double* vgrad;
    int64_t goff0;
    double* cw;
    double* cN;
    double* cR;
    uint8_t* cnodes;
    uint8_t* cnh;
    int32_t* claw;
};

// X: coordinates of the element's slots (regular nodes, then enrichment
// nodes); poly: n-gon vertex tables.
PGF_HD inline int cut_tables(int dim, int kind, const CutElem& ce, const double (*X)[3], double tet_weight_scale,
                             const geom::PolygonTable& poly, const TableOut& o) {
    const int nt = 4 + ce.n_enrich;
    int iv = 0;
    int64_t ig = 0;
    for (int ie = 0; ie < ce.n_sub; ++ie) {
        const int cn = ce.sub_nn[ie];
        double CX[8][3];
        for (int k = 0; k < cn; ++k)
            for (int d = 0; d < 3; ++d) CX[k][d] = X[ce.sub_nodes[ie][k]][d];
        const int nqp = sub_nqp(dim, cn);
        double wsum = 0.0;
        double grad[MAX_EL_NODES][3];
        for (int qp = 0; qp < nqp; ++qp) {
            double xi[3] = {0, 0, 0}, W;
            double cN[8], cdN[8][3], cdx[8][3];
            if (dim == 3) { geom::tet4_qp(qp, xi, W); geom::tet4_shape(xi, cN, cdN); }
            else if (cn == 3) { geom::tri3_qp(qp, xi, W); geom::tri3_shape(xi, cN, cdN); }
            else if (cn == 4) { geom::quad4_qp(qp, xi, W); geom::quad4_shape(xi, cN, cdN); }
            else { geom::polygon_qp(cn, poly.v[cn], qp, xi, W); geom::polygon_shape(cn, poly.v[cn], xi, cN, cdN); }
            double Jc = geom::transform_gradients(dim, cn, CX, cdN, cdx);
            if (Jc < 0.0) return TAB_NEG_JACOBIAN;
            double gpt[3] = {0, 0, 0};
            for (int d = 0; d < dim; ++d)
                for (int k = 0; k < cn; ++k) gpt[d] += CX[k][d] * cN[k];
            double pxi[3] = {0, 0, 0};
            parent_inverse_map(kind, X, gpt, pxi);
            if (!parent_inside(kind, pxi)) return TAB_OUTSIDE;
            double pN[4], pdN[4][3], pdx[4][3];
            if (kind == QUAD4) geom::quad4_shape(pxi, pN, pdN);
            else geom::tet4_shape(pxi, pN, pdN);
            geom::transform_gradients(dim, 4, X, pdN, pdx);
            double chk = 0.0;
            for (int d = 0; d < dim; ++d) {
                double s = 0.0;
                for (int n = 0; n < 4; ++n) s += X[n][d] * pN[n];
                chk += (s - gpt[d]) * (s - gpt[d]);
            }
            if (sqrt(chk) > 1e-9) return TAB_MAP_CHECK;
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
            o.vw[iv] = w;
            o.vmat[iv] = ce.sub_mat[ie];
            o.vgoff[iv] = o.goff0 + ig;
            ++iv;
            for (int s = 0; s < nt; ++s)
                for (int d = 0; d < dim; ++d) o.vgrad[ig++] = grad[s][d];
        }
        if (dim == 3) {
            o.vw[iv] = wsum * tet_weight_scale;
            o.vmat[iv] = ce.sub_mat[ie];
            o.vgoff[iv] = o.goff0 + ig;
            ++iv;
            for (int s = 0; s < nt; ++s)
                for (int d = 0; d < dim; ++d) o.vgrad[ig++] = grad[s][d];
        }
    }
    int ic = 0;
    for (int c = 0; c < ce.n_coh; ++c) {
        const int nh = ce.coh_nn[c] / 2;
        const int sd = dim - 1;  // surface dimension
        int nqp = dim == 2 ? geom::kEdge2NQP : (nh == 3 ? geom::kTri3NQP : geom::kQuad4NQP);
        for (int qp = 0; qp < nqp; ++qp, ++ic) {
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
            dA = sqrt(dA);
            double R[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int i = 0; i < dim; ++i) R[(dim - 1) * dim + i] = nrm[i] / dA;
            for (int a = 0; a < sd; ++a) {
                double mag = 0.0;
                for (int i = 0; i < dim; ++i) mag += Jt[a][i] * Jt[a][i];
                mag = sqrt(mag);
                for (int i = 0; i < dim; ++i) R[a * dim + i] = Jt[a][i] / mag;
            }
            o.cw[ic] = dA * W;
            for (int n = 0; n < 4; ++n) o.cN[ic * 4 + n] = n < nh ? N[n] : 0.0;
            for (int k = 0; k < 9; ++k) o.cR[ic * 9 + k] = R[k];
            for (int k = 0; k < 8; ++k) o.cnodes[ic * 8 + k] = (uint8_t)(k < 2 * nh ? ce.coh_nodes[c][k] : 0);
            o.cnh[ic] = (uint8_t)nh;
            o.claw[ic] = ce.coh_law[c];
        }
    }
    return TAB_OK;
}

}  // namespace igfem
}  // namespace pgf
