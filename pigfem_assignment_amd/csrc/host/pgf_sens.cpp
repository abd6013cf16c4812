// Shape-sensitivity tables of cut elements with respect to inclusion radii.
//
// Mirrors SensitivityShapeParameter::compute_velocity
// (src/SensitivityShapeParameter.cpp:49-182) and Elem::ShapeFunctionPartials /
// ChildShapeFunctionPartials (src/elem.cpp:141-292):
//   * every enrichment node moves with the radius of its inclusion along its
//     parent edge, v = (p0 - p1) r / ((p0 - p1) . (x - c))
//     (Circle_Inclusion / Sphere_Inclusion::getIntersectionVelocity,
//     src/InclusionCircle.cpp:176-196, src/InclusionSphere.cpp:191-215);
//   * at every volumetric integration point of a cut element: the velocity
//     divergence div v of the sub-element's interpolated velocity field, the
//     derivative of the enrichment-slot gradients, -J^-1 dJ J^-1 dN/dxi with
//     dJ = sum_n dN_n/dxi (x) v_n (child map), and the derivative of the
//     regular-slot gradients at the moving point (parent map).
// The parent-slot term: the exact material derivative
//   d(grad N) = J^-1 (H w - dJ grad N),  w = J^-T v,  dJ = sum_n (H_n w) (x) X_n
// with H the parent-coordinate Hessian (zero for linear tets), or, with
// `as_reference`, the reference's expression, which uses J^-1 v for w and
// keeps only the last parent node's term of dJ, evaluated with J v
// (src/elem.cpp:186-221: the loop assigns instead of accumulating).
// Points are laid out exactly as the volumetric tables of store_shape_functions
// (3D sub-element points merged into one).
#include <cmath>
#include <sstream>
#include <vector>

#include "pgf_geom.hpp"
#include "pgf_mesh.hpp"
#include "pgf_sens.hpp"

namespace pgf {

namespace {

struct Jac {
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};    // J[a][i] = sum_n dN_n/dxi_a x_n[i]
    double inv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // inv[i][a]
    double det = 0.0;
};

Jac jacobian(int dim, int nn, const double (*X)[3], const double (*dN)[3]) {
    Jac r;
    for (int a = 0; a < dim; ++a)
        for (int i = 0; i < dim; ++i)
            for (int n = 0; n < nn; ++n) r.J[a][i] += dN[n][a] * X[n][i];
    const auto& J = r.J;
    if (dim == 2) {
        r.det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        r.inv[0][0] = J[1][1] / r.det; r.inv[0][1] = -J[0][1] / r.det;
        r.inv[1][0] = -J[1][0] / r.det; r.inv[1][1] = J[0][0] / r.det;
    } else {
        r.det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        for (int i = 0; i < 3; ++i)
            for (int a = 0; a < 3; ++a) {
                const int a0 = (i + 1) % 3, a1 = (i + 2) % 3, b0 = (a + 1) % 3, b1 = (a + 2) % 3;
                r.inv[i][a] = (J[b0][a0] * J[b1][a1] - J[b0][a1] * J[b1][a0]) / r.det;
            }
    }
    return r;
}

// parent point of a global point: affine for tets, Newton for bilinear quads
void inverse_map(ElemKind kind, const double (*P)[3], const double* g, double* xi) {
    const int dim = kind == TET4 ? 3 : 2;
    xi[0] = kind == TET4 ? 0.25 : 0.0;
    xi[1] = xi[0];
    xi[2] = kind == TET4 ? 0.25 : 0.0;
    for (int it = 0; it < 50; ++it) {
        double N[4], dN[4][3];
        if (kind == TET4) geom::tet4_shape(xi, N, dN);
        else geom::quad4_shape(xi, N, dN);
        double f[3] = {0, 0, 0};
        for (int n = 0; n < 4; ++n)
            for (int i = 0; i < dim; ++i) f[i] += N[n] * P[n][i];
        for (int i = 0; i < dim; ++i) f[i] -= g[i];
        const Jac j = jacobian(dim, 4, P, dN);
        double s = 0.0;
        for (int a = 0; a < dim; ++a) {
            // dxi_a = sum_i (J^-T)_{a i} f_i = sum_i inv[i][a] f_i
            double d = 0.0;
            for (int i = 0; i < dim; ++i) d += j.inv[i][a] * f[i];
            xi[a] -= d;
            s += std::fabs(d);
        }
        if (s < 1e-16) break;
    }
}

// parent-coordinate Hessians of the bilinear quad (only the mixed derivative)
void quad4_hessian(double (*H)[3][3]) {
    const double m[4] = {0.25, -0.25, 0.25, -0.25};
    for (int n = 0; n < 4; ++n) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) H[n][a][b] = 0.0;
        H[n][0][1] = H[n][1][0] = m[n];
    }
}

}  // namespace

void build_shape_sensitivity(const Mesh& m, bool as_reference, ShapeSensTables& t) {
    const int dim = m.dim();
    const int64_t nn = m.n_nodes();
    const int64_t ne = m.n_enrich_nodes();
    const auto& edges = m.enrich_edge();
    const auto& incl = m.inclusions();
    t.vel.assign((size_t)3 * ne, 0.0);
    for (int64_t k = 0; k < ne; ++k) {
        const auto& ed = edges[k];
        const Inclusion& inc = incl[ed[2]];
        double al[3] = {0, 0, 0}, dot = 0.0;
        for (int d = 0; d < dim; ++d) {
            al[d] = m.X(d)[ed[0]] - m.X(d)[ed[1]];
            dot += al[d] * (m.X(d)[nn + k] - inc.c[d]);
        }
        const double val = inc.r / dot;
        for (int d = 0; d < dim; ++d) t.vel[3 * k + d] = al[d] * val;
    }
    t.dgrad.assign(m.vpt_grad.size(), 0.0);
    t.divv.assign(m.vpt_w.size(), 0.0);
    t.param.assign(m.cut().size(), -1);
    int64_t pt = 0;
    double H[4][3][3];
    quad4_hessian(H);
    for (size_t ci = 0; ci < m.cut().size(); ++ci) {
        const CutElem& ce = m.cut()[ci];
        const int64_t e = ce.elem;
        const int nt = 4 + ce.n_enrich;
        t.param[ci] = ce.enrich_incl[0];
        double X[MAX_EL_NODES][3], V[MAX_EL_NODES][3];
        for (int s = 0; s < nt; ++s) {
            const int32_t g = m.elem_node(e, s);
            for (int d = 0; d < 3; ++d) {
                X[s][d] = m.X(d)[g];
                V[s][d] = s < 4 ? 0.0 : t.vel[3 * (g - nn) + d];
            }
        }
        for (int ie = 0; ie < ce.n_sub; ++ie) {
            const int cn = ce.sub_nn[ie];
            double CX[8][3], CV[8][3];
            for (int k = 0; k < cn; ++k)
                for (int d = 0; d < 3; ++d) {
                    CX[k][d] = X[ce.sub_nodes[ie][k]][d];
                    CV[k][d] = V[ce.sub_nodes[ie][k]][d];
                }
            int nqp;
            if (dim == 3) nqp = 1;  // linear tet: one merged point, every quantity constant
            else if (cn == 3) nqp = geom::kTri3NQP;
            else if (cn == 4) nqp = geom::kQuad4NQP;
            else nqp = cn * geom::kTri3NQP;
            for (int qp = 0; qp < nqp; ++qp, ++pt) {
                if (pt >= (int64_t)m.vpt_w.size()) throw Error(PGF_ERR_STATE, "sensitivity points out of step");
                double xi[3] = {0, 0, 0}, W, cN[8], cdN[8][3];
                if (dim == 3) { geom::tet4_qp(0, xi, W); geom::tet4_shape(xi, cN, cdN); }
                else if (cn == 3) { geom::tri3_qp(qp, xi, W); geom::tri3_shape(xi, cN, cdN); }
                else if (cn == 4) { geom::quad4_qp(qp, xi, W); geom::quad4_shape(xi, cN, cdN); }
                else { geom::polygon_qp(cn, qp, xi, W); geom::polygon_shape(cn, xi, cN, cdN); }
                const Jac cj = jacobian(dim, cn, CX, cdN);
                double dJ[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                for (int a = 0; a < dim; ++a)
                    for (int i = 0; i < dim; ++i)
                        for (int n = 0; n < cn; ++n) dJ[a][i] += cdN[n][a] * CV[n][i];
                double divv = 0.0, gx[3] = {0, 0, 0}, vx[3] = {0, 0, 0};
                double* dg = &t.dgrad[m.vpt_grad_off[pt]];
                for (int n = 0; n < cn; ++n) {
                    double cdx[3] = {0, 0, 0};
                    for (int i = 0; i < dim; ++i)
                        for (int a = 0; a < dim; ++a) cdx[i] += cj.inv[i][a] * cdN[n][a];
                    for (int i = 0; i < dim; ++i) divv += cdx[i] * CV[n][i];
                    for (int d = 0; d < dim; ++d) {
                        gx[d] += cN[n] * CX[n][d];
                        vx[d] += cN[n] * CV[n][d];
                    }
                    const int slot = ce.sub_nodes[ie][n];
                    if (slot < 4) continue;
                    // -J^-1 (dJ (J^-1 dN))
                    double tmp[3] = {0, 0, 0};
                    for (int a = 0; a < dim; ++a)
                        for (int i = 0; i < dim; ++i) tmp[a] += dJ[a][i] * cdx[i];
                    for (int i = 0; i < dim; ++i) {
                        double s = 0.0;
                        for (int a = 0; a < dim; ++a) s += cj.inv[i][a] * tmp[a];
                        dg[slot * dim + i] = -s;
                    }
                }
                t.divv[pt] = divv;
                if (dim == 3) continue;  // linear parent: its gradients do not change with the point
                double pxi[3] = {0, 0, 0}, pN[4], pdN[4][3];
                inverse_map(m.kind(), X, gx, pxi);
                geom::quad4_shape(pxi, pN, pdN);
                const Jac pj = jacobian(2, 4, X, pdN);
                double w[2], dJp[2][3] = {{0, 0, 0}, {0, 0, 0}};
                if (!as_reference) {
                    for (int a = 0; a < 2; ++a) w[a] = pj.inv[0][a] * vx[0] + pj.inv[1][a] * vx[1];
                    for (int n = 0; n < 4; ++n)
                        for (int a = 0; a < 2; ++a) {
                            const double hw = H[n][a][0] * w[0] + H[n][a][1] * w[1];
                            for (int i = 0; i < 2; ++i) dJp[a][i] += hw * X[n][i];
                        }
                } else {
                    for (int i = 0; i < 2; ++i) w[i] = pj.inv[i][0] * vx[0] + pj.inv[i][1] * vx[1];
                    const double Jv[2] = {pj.J[0][0] * vx[0] + pj.J[0][1] * vx[1], pj.J[1][0] * vx[0] + pj.J[1][1] * vx[1]};
                    for (int a = 0; a < 2; ++a) {
                        const double ret = H[3][a][0] * Jv[0] + H[3][a][1] * Jv[1];
                        for (int i = 0; i < 2; ++i) dJp[a][i] = ret * X[3][i];
                    }
                }
                for (int n = 0; n < 4; ++n) {
                    double pdx[2] = {0, 0};
                    for (int i = 0; i < 2; ++i)
                        for (int a = 0; a < 2; ++a) pdx[i] += pj.inv[i][a] * pdN[n][a];
                    double r[2];
                    for (int a = 0; a < 2; ++a)
                        r[a] = (H[n][a][0] * w[0] + H[n][a][1] * w[1]) - (dJp[a][0] * pdx[0] + dJp[a][1] * pdx[1]);
                    for (int i = 0; i < 2; ++i) dg[n * 2 + i] = pj.inv[i][0] * r[0] + pj.inv[i][1] * r[1];
                }
            }
        }
        if (pt != m.vpt_ptr[ci + 1]) throw Error(PGF_ERR_STATE, "sensitivity points out of step");
    }
}

}  // namespace pgf
