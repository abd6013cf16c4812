// Quadrature rules and shape functions used to build the per-quadrature-point
// tables of cut (IGFEM) elements on the host.
//
// Rules are the reference's default order-3 rules:
//   * EDGE2 / COHEDGE2: 2-point Gauss, xi = -+1/sqrt(3), w = 1
//     (src/edge2.cpp:111-118, src/CohEdge2.cpp:70-77)
//   * QUAD4 / COHQUAD4: tensor product of the edge rule, qp -> (qp%2, qp/2)
//     (src/quad4.cpp:97-115, src/CohQuad4.cpp:62-80)
//   * TRI3 / COHTRI3: 4-point rule with the -27/96 centroid weight, weights sum
//     to 1/2 (src/tri3.cpp:121-127, src/CohTri3.cpp:79-85)
//   * TET4: 5-point rule with the -0.8 centroid weight, weights sum to 1
//     (src/tet4.cpp:114-121).  The tet Jacobian is 6*volume, so a tet volume
//     integral comes out 6x the volume; that is the reference's convention and
//     is kept (a tet_weight_scale of 1/6 corrects it when asked for).
//   * POLYGON: fan of n triangles of the regular reference n-gon, each with the
//     TRI3 rule (src/Polygon.cpp:51-126).
// Shape functions: src/quad4.cpp:129-149, src/tri3.cpp:191-207,
// src/tet4.cpp:163-182, Wachspress polygon src/Polygon.cpp:128-227,
// cohesive src/CohEdge2.cpp:186-210, src/CohTri3.cpp:170-205,
// src/CohQuad4.cpp:82-124; Jacobian/gradient transform src/elem.cpp:112-136.
#pragma once

#include <cmath>
#include "pgf_common.hpp"

namespace pgf {
namespace geom {

constexpr double kPI = 3.14159265358979323846;
constexpr double kGauss2 = 0.57735026918962576450914878050196;

// ---- reference rules ------------------------------------------------------
PGF_HD inline void edge2_qp(int qp, double& xi, double& w) {
    xi = qp == 0 ? -kGauss2 : kGauss2;
    w = 1.0;
}
PGF_HD inline void quad4_qp(int qp, double* xi, double& w) {
    double w0, w1;
    edge2_qp(qp % 2, xi[0], w0);
    edge2_qp(qp / 2, xi[1], w1);
    w = w0 * w1;
}
PGF_HD inline void tri3_qp(int qp, double* xi, double& w) {
    constexpr double P[4][2] = {{0.333333333333333333333333333333333, 0.333333333333333333333333333333333},
                                {0.2, 0.2}, {0.2, 0.6}, {0.6, 0.2}};
    constexpr double W[4] = {-0.28125, 0.260416666666666666666666666666665,
                             0.260416666666666666666666666666665, 0.260416666666666666666666666666665};
    xi[0] = P[qp][0]; xi[1] = P[qp][1]; w = W[qp];
}
PGF_HD inline void tet4_qp(int qp, double* xi, double& w) {
    constexpr double P[5][3] = {{0.25, 0.25, 0.25},
                                {0.5, 0.166666666666667, 0.166666666666667},
                                {0.166666666666667, 0.166666666666667, 0.166666666666667},
                                {0.166666666666667, 0.166666666666667, 0.5},
                                {0.166666666666667, 0.5, 0.166666666666667}};
    constexpr double W[5] = {-0.8, 0.45, 0.45, 0.45, 0.45};
    xi[0] = P[qp][0]; xi[1] = P[qp][1]; xi[2] = P[qp][2]; w = W[qp];
}
constexpr int kQuad4NQP = 4, kTri3NQP = 4, kTet4NQP = 5, kEdge2NQP = 2;

// ---- shape functions (parent coordinates) -----------------------------------
PGF_HD inline void quad4_shape(const double* xi, double* N, double (*dN)[3]) {
    double r = xi[0], s = xi[1];
    N[0] = 0.25 * (1.0 - r) * (1.0 - s); N[1] = 0.25 * (1.0 + r) * (1.0 - s);
    N[2] = 0.25 * (1.0 + r) * (1.0 + s); N[3] = 0.25 * (1.0 - r) * (1.0 + s);
    if (dN) {
        dN[0][0] = -0.25 * (1.0 - s); dN[0][1] = -0.25 * (1.0 - r);
        dN[1][0] = 0.25 * (1.0 - s);  dN[1][1] = -0.25 * (1.0 + r);
        dN[2][0] = 0.25 * (1.0 + s);  dN[2][1] = 0.25 * (1.0 + r);
        dN[3][0] = -0.25 * (1.0 + s); dN[3][1] = 0.25 * (1.0 - r);
    }
}
PGF_HD inline void tri3_shape(const double* xi, double* N, double (*dN)[3]) {
    N[0] = 1.0 - xi[0] - xi[1]; N[1] = xi[0]; N[2] = xi[1];
    if (dN) {
        dN[0][0] = -1.0; dN[0][1] = -1.0;
        dN[1][0] = 1.0;  dN[1][1] = 0.0;
        dN[2][0] = 0.0;  dN[2][1] = 1.0;
    }
}
PGF_HD inline void tet4_shape(const double* xi, double* N, double (*dN)[3]) {
    N[0] = 1.0 - xi[0] - xi[1] - xi[2]; N[1] = xi[0]; N[2] = xi[1]; N[3] = xi[2];
    if (dN) {
        dN[0][0] = -1.0; dN[0][1] = -1.0; dN[0][2] = -1.0;
        dN[1][0] = 1.0;  dN[1][1] = 0.0;  dN[1][2] = 0.0;
        dN[2][0] = 0.0;  dN[2][1] = 1.0;  dN[2][2] = 0.0;
        dN[3][0] = 0.0;  dN[3][1] = 0.0;  dN[3][2] = 1.0;
    }
}

// Wachspress coordinates on the regular reference n-gon whose vertex k sits at
// angle 2*pi*(k+1)/n on the unit circle; alpha_k = 1/(A_{k-1} A_k) with A_k
// the signed area of triangle (x, p_k, p_{k+1}).  The vertices come from a
// table V[k] = polygon_vertex(nn, k) built on the host, so host and device
// use the same bits.
inline void polygon_vertex(int nn, int k, double* p) {
    p[0] = std::cos(2.0 * kPI * (k + 1) / nn);
    p[1] = std::sin(2.0 * kPI * (k + 1) / nn);
}
// Vertex tables of the n-gons the sub-elements use (n = 0..MAX_EL_NODES;
// row n holds n vertices).
struct PolygonTable {
    double v[MAX_EL_NODES + 1][MAX_EL_NODES][2];
};
inline const PolygonTable& polygon_table() {
    static const PolygonTable t = [] {
        PolygonTable p{};
        for (int n = 3; n <= MAX_EL_NODES; ++n)
            for (int k = 0; k < n; ++k) polygon_vertex(n, k, p.v[n][k]);
        return p;
    }();
    return t;
}
PGF_HD inline void polygon_shape(int nn, const double (*V)[2], const double* x, double* N, double (*dN)[3]) {
    double A[MAX_EL_NODES + 1], dA[MAX_EL_NODES + 1][2];
    for (int n = 0; n < nn; ++n) {
        double p0[2] = {x[0], x[1]}, p1[2], p2[2];
        p1[0] = V[n][0]; p1[1] = V[n][1];
        p2[0] = V[(n + 1) % nn][0]; p2[1] = V[(n + 1) % nn][1];
        A[n + 1] = 0.5 * ((p1[0] * p2[1] - p1[1] * p2[0]) - (p0[0] * p2[1] - p0[1] * p2[0]) +
                          (p0[0] * p1[1] - p0[1] * p1[0]));
        dA[n + 1][0] = 0.5 * (p2[1] - p1[1]);
        dA[n + 1][1] = 0.5 * (p1[0] - p2[0]);
    }
    A[0] = A[nn]; dA[0][0] = dA[nn][0]; dA[0][1] = dA[nn][1];
    double alpha[MAX_EL_NODES], dal[MAX_EL_NODES][2];
    double sa = 0.0, sd0 = 0.0, sd1 = 0.0;
    for (int n = 0; n < nn; ++n) {
        alpha[n] = 1.0 / (A[n] * A[n + 1]);
        dal[n][0] = -1.0 * alpha[n] * (dA[n][0] / A[n] + dA[n + 1][0] / A[n + 1]);
        dal[n][1] = -1.0 * alpha[n] * (dA[n][1] / A[n] + dA[n + 1][1] / A[n + 1]);
        sa += alpha[n]; sd0 += dal[n][0]; sd1 += dal[n][1];
    }
    for (int n = 0; n < nn; ++n) {
        double Nn = alpha[n] / sa;
        N[n] = Nn;
        if (dN) {
            dN[n][0] = (dal[n][0] - Nn * sd0) / sa;
            dN[n][1] = (dal[n][1] - Nn * sd1) / sa;
        }
    }
}
inline void polygon_shape(int nn, const double* x, double* N, double (*dN)[3]) {
    polygon_shape(nn, polygon_table().v[nn], x, N, dN);
}
// Fan quadrature point qp of the regular n-gon: triangle qp/4 = (origin, p_k,
// p_{k+1}) with the 4-point TRI3 rule (src/Polygon.cpp:82-126).
PGF_HD inline void polygon_qp(int nn, const double (*V)[2], int qp, double* xi, double& w) {
    double txi[2], tw;
    tri3_qp(qp % kTri3NQP, txi, tw);
    double tN[3], tdN[3][3];
    tri3_shape(txi, tN, tdN);
    int t = qp / kTri3NQP;
    double pT[3][2] = {{0.0, 0.0}, {V[t][0], V[t][1]}, {V[(t + 1) % nn][0], V[(t + 1) % nn][1]}};
    double J0[2][2] = {{0, 0}, {0, 0}};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int n = 0; n < 3; ++n) J0[i][j] += pT[n][i] * tdN[n][j];
    w = (J0[0][0] * J0[1][1] - J0[0][1] * J0[1][0]) * tw;
    xi[0] = 0.0; xi[1] = 0.0;
    for (int n = 0; n < 3; ++n) { xi[0] += tN[n] * pT[n][0]; xi[1] += tN[n] * pT[n][1]; }
}
inline void polygon_qp(int nn, int qp, double* xi, double& w) { polygon_qp(nn, polygon_table().v[nn], qp, xi, w); }

// Elem::transform_gradients (src/elem.cpp:123-136): Jac(a,i) = sum_n dNdr[n][a] x_n[i],
// J = det(Jac), dNdx[n] = Jac^{-1} dNdr[n].  Returns J.
PGF_HD inline double transform_gradients(int dim, int nn, const double (*X)[3], const double (*dNdr)[3],
                                  double (*dNdx)[3]) {
    double Jc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < dim; ++a)
        for (int i = 0; i < dim; ++i)
            for (int n = 0; n < nn; ++n) Jc[a][i] += dNdr[n][a] * X[n][i];
    double det, inv[3][3];
    if (dim == 2) {
        det = Jc[0][0] * Jc[1][1] - Jc[0][1] * Jc[1][0];
        inv[0][0] = Jc[1][1] / det; inv[0][1] = -Jc[0][1] / det;
        inv[1][0] = -Jc[1][0] / det; inv[1][1] = Jc[0][0] / det;
    } else {
        det = Jc[0][0] * (Jc[1][1] * Jc[2][2] - Jc[1][2] * Jc[2][1]) -
              Jc[0][1] * (Jc[1][0] * Jc[2][2] - Jc[1][2] * Jc[2][0]) +
              Jc[0][2] * (Jc[1][0] * Jc[2][1] - Jc[1][1] * Jc[2][0]);
        inv[0][0] = (Jc[1][1] * Jc[2][2] - Jc[1][2] * Jc[2][1]) / det;
        inv[0][1] = (Jc[0][2] * Jc[2][1] - Jc[0][1] * Jc[2][2]) / det;
        inv[0][2] = (Jc[0][1] * Jc[1][2] - Jc[0][2] * Jc[1][1]) / det;
        inv[1][0] = (Jc[1][2] * Jc[2][0] - Jc[1][0] * Jc[2][2]) / det;
        inv[1][1] = (Jc[0][0] * Jc[2][2] - Jc[0][2] * Jc[2][0]) / det;
        inv[1][2] = (Jc[0][2] * Jc[1][0] - Jc[0][0] * Jc[1][2]) / det;
        inv[2][0] = (Jc[1][0] * Jc[2][1] - Jc[1][1] * Jc[2][0]) / det;
        inv[2][1] = (Jc[0][1] * Jc[2][0] - Jc[0][0] * Jc[2][1]) / det;
        inv[2][2] = (Jc[0][0] * Jc[1][1] - Jc[0][1] * Jc[1][0]) / det;
    }
    for (int n = 0; n < nn; ++n)
        for (int i = 0; i < dim; ++i) {
            double s = 0.0;
            for (int a = 0; a < dim; ++a) s += inv[i][a] * dNdr[n][a];
            dNdx[n][i] = s;
        }
    return det;
}

}  // namespace geom
}  // namespace pgf
