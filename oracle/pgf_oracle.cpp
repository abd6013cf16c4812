// CPU oracle for the PIGFEM assemble-and-solve path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates the reference's algorithm on
// the CPU so the HIP path can be checked against it; only tests/, the smoke
// check and the benchmark's CPU-baseline leg load it.  It is never linked into
// or called by libpgfem.so.
//
// What it restates (paths relative to PIGFEM4.6.1/):
//   * QP tables of every element: Mesh::store_shape_functions
//     (src/Mesh.cpp:1443-1605) with Elem::ShapeFunctions / transform_gradients
//     (src/elem.cpp:112-136), quadrature rules of src/quad4.cpp:65-115,
//     src/tri3.cpp:121-127, src/tet4.cpp:114-121, src/Polygon.cpp:51-227,
//     cohesive shapes and rotation src/CohesiveElem.cpp:130-242.  Every tet QP
//     is kept separately (no merging).
//   * Dof numbering: DofObject::distribute_dofs (src/DofObject.cpp:241-538).
//   * Sparsity: Solver::populate_local_neighbors (src/Solver.cpp:723-829) as
//     std::set per node, split into K_ff/K_fp/K_pf/K_pp.
//   * Element matrices: Assembler::assemble_elem_linear / assemble_elem_nonlinear
//     (src/Assembler.cpp:94-207, 305-553) with the dN-variant SmallStrainKFast
//     (src/Utilities.cpp:196-296) or the B-matrix variant (:428-450, dgemm
//     order) selected by `bvariant`, cohesive KernelCohesive
//     (src/AssemblerSmallStrainNonlinear.cpp:73-91, src/Utilities.cpp:465-531),
//     OP-exponential / bilinear laws (src/Material_OPCohesive.cpp:38-319).
//   * Scatter: Assembler::set_matrix / set_vec (src/Assembler.cpp:811-879)
//     into std::map rows (MatSetValue ADD_VALUES).
//   * Solve: Jacobi-PCG in FP64 to a tight tolerance (the reference's direct
//     MUMPS solve, src/Solver.cpp:71-79, has no offline equivalent); Newton
//     SolverNonlinearKSP::nonlinearStep (src/SolverNonlinearKSP.cpp:55-159) and
//     the load-step loop of src/SolverNonlinear.cpp:111-397.
//   * Node detection (src/Mesh.cpp:2838-2918) and the enrichment-node set
//     (src/Mesh.cpp:3064-3134) for the integer flag checks, with the
//     intersection taken as the closed-form root of the level-set quadratic.
// Inputs: geometry and the cut-element topology (sub-element and cohesive node
// lists) exported by the host library; the topology itself is pinned by the
// tiling and material property tests, not by this file.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#define ORC_API extern "C" __attribute__((visibility("default")))

namespace {

using Vec = std::vector<double>;
using Mat = std::vector<std::vector<double>>;

Mat zeros(int r, int c) { return Mat(r, Vec(c, 0.0)); }

double det2(const Mat& J) { return J[0][0] * J[1][1] - J[0][1] * J[1][0]; }
double det3(const Mat& J) {
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}
Mat inverse(const Mat& J) {
    int n = (int)J.size();
    Mat I = zeros(n, n);
    if (n == 2) {
        double d = det2(J);
        I[0][0] = J[1][1] / d; I[0][1] = -J[0][1] / d; I[1][0] = -J[1][0] / d; I[1][1] = J[0][0] / d;
    } else {
        double d = det3(J);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                int a0 = (j + 1) % 3, a1 = (j + 2) % 3, b0 = (i + 1) % 3, b1 = (i + 2) % 3;
                I[i][j] = (J[a0][b0] * J[a1][b1] - J[a0][b1] * J[a1][b0]) / d;
            }
    }
    return I;
}

// ---------------- element shape functions (parent coordinates) -------------
const double G2 = 0.57735026918962576450914878050196;
struct Rule { std::vector<Vec> xi; Vec w; };
Rule rule_quad() { Rule r; for (int q = 0; q < 4; ++q) { r.xi.push_back({q % 2 ? G2 : -G2, q / 2 ? G2 : -G2}); r.w.push_back(1.0); } return r; }
Rule rule_tri() {
    Rule r;
    r.xi = {{1.0 / 3.0, 1.0 / 3.0}, {0.2, 0.2}, {0.2, 0.6}, {0.6, 0.2}};
    r.w = {-0.28125, 0.260416666666666666666666666666665, 0.260416666666666666666666666666665,
           0.260416666666666666666666666666665};
    r.xi[0] = {0.333333333333333333333333333333333, 0.333333333333333333333333333333333};
    return r;
}
Rule rule_tet() {
    Rule r;
    const double a = 0.166666666666667;
    r.xi = {{0.25, 0.25, 0.25}, {0.5, a, a}, {a, a, a}, {a, a, 0.5}, {a, 0.5, a}};
    r.w = {-0.8, 0.45, 0.45, 0.45, 0.45};
    return r;
}
Rule rule_edge() { Rule r; r.xi = {{-G2}, {G2}}; r.w = {1.0, 1.0}; return r; }

// n-gon (Wachspress on the regular reference polygon, vertex k at angle 2 pi (k+1)/n)
void poly_vertex(int nn, int k, double* p) {
    const double PI = 3.14159265358979323846;
    p[0] = std::cos(2.0 * PI * (k + 1) / nn);
    p[1] = std::sin(2.0 * PI * (k + 1) / nn);
}
double tri_area(const double* a, const double* b, const double* c) {
    return 0.5 * ((b[0] * c[1] - b[1] * c[0]) - (a[0] * c[1] - a[1] * c[0]) + (a[0] * b[1] - a[1] * b[0]));
}
void poly_shape(int nn, const Vec& x, Vec& N, Mat& dN) {
    Vec A(nn + 1);
    Mat dA = zeros(nn + 1, 2);
    for (int n = 0; n < nn; ++n) {
        double p1[2], p2[2];
        poly_vertex(nn, n, p1);
        poly_vertex(nn, (n + 1) % nn, p2);
        A[n + 1] = tri_area(x.data(), p1, p2);
        dA[n + 1][0] = 0.5 * (p2[1] - p1[1]);
        dA[n + 1][1] = 0.5 * (p1[0] - p2[0]);
    }
    A[0] = A[nn];
    dA[0] = dA[nn];
    Vec al(nn);
    Mat dal = zeros(nn, 2);
    double sa = 0, sd[2] = {0, 0};
    for (int n = 0; n < nn; ++n) {
        al[n] = 1.0 / (A[n] * A[n + 1]);
        for (int i = 0; i < 2; ++i) dal[n][i] = -1.0 * al[n] * (dA[n][i] / A[n] + dA[n + 1][i] / A[n + 1]);
        sa += al[n];
        sd[0] += dal[n][0];
        sd[1] += dal[n][1];
    }
    N.assign(nn, 0.0);
    dN = zeros(nn, 2);
    for (int n = 0; n < nn; ++n) {
        N[n] = al[n] / sa;
        for (int i = 0; i < 2; ++i) dN[n][i] = (dal[n][i] - N[n] * sd[i]) / sa;
    }
}
Rule rule_poly(int nn) {
    Rule tri = rule_tri(), r;
    for (int t = 0; t < nn; ++t)
        for (int q = 0; q < 4; ++q) {
            double o[2] = {0, 0}, p1[2], p2[2];
            poly_vertex(nn, t, p1);
            poly_vertex(nn, (t + 1) % nn, p2);
            double s = tri.xi[q][0], u = tri.xi[q][1];
            double N0 = 1 - s - u;
            r.xi.push_back({N0 * o[0] + s * p1[0] + u * p2[0], N0 * o[1] + s * p1[1] + u * p2[1]});
            // J0 = sum_n pT[n][i] dN[n][j]; det
            double J00 = -o[0] + p1[0], J01 = -o[0] + p2[0], J10 = -o[1] + p1[1], J11 = -o[1] + p2[1];
            r.w.push_back((J00 * J11 - J01 * J10) * tri.w[q]);
        }
    return r;
}
// shape functions and reference gradients of a volumetric element with nn nodes
void shape(int dim, int nn, const Vec& xi, Vec& N, Mat& dN) {
    if (dim == 3) {
        N = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
        dN = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    } else if (nn == 3) {
        N = {1.0 - xi[0] - xi[1], xi[0], xi[1]};
        dN = {{-1, -1}, {1, 0}, {0, 1}};
    } else if (nn == 4) {
        double r = xi[0], s = xi[1];
        N = {0.25 * (1 - r) * (1 - s), 0.25 * (1 + r) * (1 - s), 0.25 * (1 + r) * (1 + s), 0.25 * (1 - r) * (1 + s)};
        dN = {{-0.25 * (1 - s), -0.25 * (1 - r)}, {0.25 * (1 - s), -0.25 * (1 + r)}, {0.25 * (1 + s), 0.25 * (1 + r)},
              {-0.25 * (1 + s), 0.25 * (1 - r)}};
    } else {
        poly_shape(nn, xi, N, dN);
    }
}
// Elem::ShapeFunctions: global gradients and J
double transform(int dim, const Mat& X, const Mat& dNdr, Mat& dNdx) {
    int nn = (int)X.size();
    Mat Jac = zeros(dim, dim);
    for (int a = 0; a < dim; ++a)
        for (int i = 0; i < dim; ++i)
            for (int n = 0; n < nn; ++n) Jac[a][i] += dNdr[n][a] * X[n][i];
    double J = dim == 2 ? det2(Jac) : det3(Jac);
    Mat Ji = inverse(Jac);
    dNdx = zeros(nn, dim);
    for (int n = 0; n < nn; ++n)
        for (int i = 0; i < dim; ++i)
            for (int a = 0; a < dim; ++a) dNdx[n][i] += Ji[i][a] * dNdr[n][a];
    return J;
}

struct Law { int type; double sc, dc, df, beta, visc, tfloor; };

struct QP {
    bool coh = false;
    double w = 0;        // J*W
    int mat = 0;         // volumetric material or cohesive law
    Mat grad;            // [n_total][dim] (volumetric)
    Vec N;               // cohesive shape values
    Mat R;               // cohesive rotation
    std::vector<int> map;  // cohesive element slots (lower then upper)
    int64_t isv = -1;
};

struct Problem {
    int dim = 2, kind = 0, plane_strain = 1;
    double tet_scale = 1.0;
    int64_t n_nodes = 0, n_all = 0, n_elem = 0;
    int64_t e_begin = 0, e_end = 0;  // elements whose tables are built and assembled
    Mat X;                                   // [n_all][3]
    std::vector<std::vector<int>> elem_nodes;  // all slots per element
    std::vector<std::vector<QP>> qps;        // per element
    std::vector<int> elem_mat;
    std::vector<Mat> D;                      // per material
    std::vector<Law> laws;
    int64_t n_isv = 0;
    // periodic sets of the generated grid (Mesh::build_periodic_nodesets);
    // set[0] is the primary node; empty without periodicity
    std::vector<std::vector<int>> psets;
    // dofs
    std::vector<std::vector<int>> dofs;      // [n_all][dim]
    int64_t n_free = 0, n_pres = 0;
    Vec pres_val;
    std::string err;
};

Mat isotropic_D(int dim, bool ps, double E, double nu) {
    int s = dim == 2 ? 3 : 6;
    Mat D = zeros(s, s);
    if (dim == 2) {
        if (ps) {
            double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
            D[0][0] = c * (1.0 - nu); D[0][1] = c * nu; D[1][0] = c * nu; D[1][1] = c * (1.0 - nu);
            D[2][2] = c * (1.0 - 2.0 * nu) / 2.0;
        } else {
            double c = E / (1.0 - nu * nu);
            D[0][0] = c; D[0][1] = c * nu; D[1][0] = c * nu; D[1][1] = c; D[2][2] = c * (1.0 - nu) / 2.0;
        }
    } else {
        double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) D[i][j] = c * (i == j ? 1.0 - nu : nu);
        for (int i = 3; i < 6; ++i) D[i][i] = c * (1.0 - 2.0 * nu) / 2;
    }
    return D;
}

// parent inverse map: Newton on x(xi) (affine elements converge in one step)
Vec inverse_map(int dim, const Mat& PX, const Vec& g) {
    Vec xi(dim, dim == 2 ? 0.0 : 0.25);
    for (int it = 0; it < 60; ++it) {
        Vec N;
        Mat dN;
        shape(dim, 4, xi, N, dN);
        Vec f(dim, 0.0);
        Mat J = zeros(dim, dim);
        for (int n = 0; n < 4; ++n)
            for (int i = 0; i < dim; ++i) {
                f[i] += N[n] * PX[n][i];
                for (int j = 0; j < dim; ++j) J[i][j] += PX[n][i] * dN[n][j];
            }
        for (int i = 0; i < dim; ++i) f[i] -= g[i];
        Mat Ji = inverse(J);
        double s = 0;
        for (int i = 0; i < dim; ++i) {
            double d = 0;
            for (int j = 0; j < dim; ++j) d += Ji[i][j] * f[j];
            xi[i] -= d;
            s += std::fabs(d);
        }
        if (s < 1e-16) break;
    }
    return xi;
}

void build_tables(Problem& P, const int32_t* cut, int64_t n_cut) {
    const int dim = P.dim;
    std::vector<int> cidx(P.n_elem, -1);
    for (int64_t c = 0; c < n_cut; ++c) cidx[cut[c * 128]] = (int)c;
    P.qps.assign(P.n_elem, {});
    int64_t isv = 0;
    // elem_nodes holds the 4 regular nodes on entry; enrichment slots are appended
    for (int64_t e = P.e_begin; e < P.e_end; ++e) {
        std::vector<int>& nodes = P.elem_nodes[e];
        if (cidx[e] < 0) {
            Mat PX(4);
            for (int n = 0; n < 4; ++n) PX[n] = {P.X[nodes[n]][0], P.X[nodes[n]][1], P.X[nodes[n]][2]};
            Rule r = dim == 2 ? rule_quad() : rule_tet();
            for (size_t q = 0; q < r.w.size(); ++q) {
                Vec N;
                Mat dN, dx;
                shape(dim, 4, r.xi[q], N, dN);
                double J = transform(dim, PX, dN, dx);
                QP p;
                p.w = J * r.w[q] * (dim == 3 ? P.tet_scale : 1.0);
                p.mat = P.elem_mat[e];
                p.grad = dx;
                P.qps[e].push_back(p);
            }
            continue;
        }
        const int32_t* rec = cut + (int64_t)cidx[e] * 128;
        const int ne = rec[100];
        for (int s = 0; s < ne; ++s) nodes.push_back((int)(P.n_nodes + rec[101 + s]));
        const int nt = 4 + ne;
        Mat EX(nt);
        for (int s = 0; s < nt; ++s) EX[s] = {P.X[nodes[s]][0], P.X[nodes[s]][1], P.X[nodes[s]][2]};
        Mat PX(EX.begin(), EX.begin() + 4);
        for (int ie = 0; ie < rec[18]; ++ie) {
            const int cn = rec[19 + ie];
            std::vector<int> sn(rec + 25 + ie * 8, rec + 25 + ie * 8 + cn);
            Mat CX(cn);
            for (int k = 0; k < cn; ++k) CX[k] = EX[sn[k]];
            Rule r = dim == 3 ? rule_tet() : (cn == 3 ? rule_tri() : (cn == 4 ? rule_quad() : rule_poly(cn)));
            for (size_t q = 0; q < r.w.size(); ++q) {
                Vec cN, pN;
                Mat cdN, cdx, pdN, pdx;
                shape(dim, cn, r.xi[q], cN, cdN);
                double Jc = transform(dim, CX, cdN, cdx);
                Vec g(dim, 0.0);
                for (int d = 0; d < dim; ++d)
                    for (int k = 0; k < cn; ++k) g[d] += CX[k][d] * cN[k];
                Vec pxi = inverse_map(dim, PX, g);
                shape(dim, 4, pxi, pN, pdN);
                transform(dim, PX, pdN, pdx);
                QP p;
                p.w = Jc * r.w[q] * (dim == 3 ? P.tet_scale : 1.0);
                p.mat = rec[73 + ie];
                p.grad = zeros(nt, dim);
                for (int n = 0; n < 4; ++n) p.grad[n] = pdx[n];
                for (int k = 0; k < cn; ++k)
                    if (sn[k] >= 4) p.grad[sn[k]] = cdx[k];
                P.qps[e].push_back(p);
            }
        }
        for (int ce = 0; ce < rec[79]; ++ce) {
            const int cnn = rec[80 + ce], nh = cnn / 2, sd = dim - 1;
            std::vector<int> map(rec + 82 + ce * 8, rec + 82 + ce * 8 + cnn);
            Rule r = dim == 2 ? rule_edge() : (nh == 3 ? rule_tri() : rule_quad());
            for (size_t q = 0; q < r.w.size(); ++q) {
                Vec N;
                Mat dN;
                if (dim == 2) { N = {0.5 * (1.0 - r.xi[q][0]), 0.5 * (1 + r.xi[q][0])}; dN = {{-0.5}, {0.5}}; }
                else shape(2, nh, r.xi[q], N, dN);
                Mat Jt = zeros(sd, dim);
                for (int n = 0; n < nh; ++n)
                    for (int a = 0; a < sd; ++a)
                        for (int i = 0; i < dim; ++i) Jt[a][i] += dN[n][a] * EX[map[n]][i];
                Vec nrm;
                if (dim == 2) nrm = {-Jt[0][1], Jt[0][0]};
                else nrm = {Jt[0][1] * Jt[1][2] - Jt[0][2] * Jt[1][1], Jt[0][2] * Jt[1][0] - Jt[0][0] * Jt[1][2],
                            Jt[0][0] * Jt[1][1] - Jt[0][1] * Jt[1][0]};
                double dA = 0;
                for (double v : nrm) dA += v * v;
                dA = std::sqrt(dA);
                QP p;
                p.coh = true;
                p.R = zeros(dim, dim);
                for (int i = 0; i < dim; ++i) p.R[dim - 1][i] = nrm[i] / dA;
                for (int a = 0; a < sd; ++a) {
                    double m = 0;
                    for (int i = 0; i < dim; ++i) m += Jt[a][i] * Jt[a][i];
                    m = std::sqrt(m);
                    for (int i = 0; i < dim; ++i) p.R[a][i] = Jt[a][i] / m;
                }
                p.w = dA * r.w[q];
                p.N = N;
                p.map = map;
                p.mat = rec[98 + ce];
                p.isv = isv++;
                P.qps[e].push_back(p);
            }
        }
    }
    P.n_isv = isv;
}

// OP exponential / bilinear law (returns traction, tangent, new delta_max, flags)
void law_eval(const Law& L, int dim, Vec& delta, double dmax, double dt, Vec& trac, Mat& Dc, double& dnew,
              int& flag) {
    double b2 = L.beta * L.beta;
    double norm = 0;
    for (double v : delta) norm += v * v;
    if (norm < 1e-40)
        for (double& v : delta) v = 1e-20;
    double deff = std::pow(delta[dim - 1], 2);
    for (int i = 0; i < dim - 1; ++i) deff += b2 * std::pow(delta[i], 2);
    deff = std::sqrt(deff);
    double T = 0, dT = 0;
    bool loading = deff >= dmax || deff < 0.0;
    auto bil = [&](double x, double& s) {
        if (x <= L.dc) { s = L.sc / L.dc; return L.sc / L.dc * x; }
        if (x <= L.df) { s = -L.sc / (L.df - L.dc); return -L.sc / (L.df - L.dc) * (x - L.dc) + L.sc; }
        s = 0;
        return 0.0;
    };
    if (loading) {
        if (L.type == 0) {
            T = L.sc * (deff / L.dc) * std::exp(1.0 - deff / L.dc) + (L.visc != 0 ? L.visc / L.sc * (deff - dmax) / dt / L.dc : 0.0);
            dT = (L.sc / L.dc) * (1.0 - deff / L.dc) * std::exp(1.0 - deff / L.dc) + (L.visc != 0 ? L.visc / L.sc / dt / L.dc : 0.0);
        } else {
            T = bil(deff, dT);
            if (T < L.tfloor && deff > L.dc) { T = L.tfloor; dT = 0; }
        }
        dnew = deff;
    } else {
        double s, tm = L.type == 0 ? L.sc * (dmax / L.dc) * std::exp(1.0 - dmax / L.dc) : bil(dmax, s);
        T = tm * (deff / dmax);
        dT = tm / dmax;
        dnew = dmax;
    }
    Vec br(dim);
    for (int i = 0; i < dim; ++i) br[i] = b2 * delta[i];
    br[dim - 1] += (1.0 - b2) * delta[dim - 1];
    trac.assign(dim, 0);
    for (int i = 0; i < dim; ++i) trac[i] = (T / deff) * br[i];
    Vec dd(dim);
    dd[dim - 1] = delta[dim - 1] / deff;
    for (int i = 0; i < dim - 1; ++i) dd[i] = b2 * delta[i] / deff;
    Dc = zeros(dim, dim);
    for (int i = 0; i < dim; ++i)
        for (int k = 0; k < dim; ++k)
            Dc[i][k] = br[i] * (dT * deff - T) * (dd[k] / std::pow(deff, 2)) +
                       (T / deff) * (b2 * (i == k) + (1.0 - b2) * (i == dim - 1) * (k == dim - 1));
    double nz = deff / L.dc;
    int zone = nz <= 0 ? 0 : (nz <= 1 ? 1 : (nz <= 5.0 ? 2 : 3));
    flag = (loading ? 1 : 0) | (zone << 1) | (dnew > L.dc ? 8 : 0) | ((L.type == 1 && dnew >= L.df) ? 16 : 0);
}

// K_qp = B^T D B, dN variant (SmallStrainKFast) or B-matrix variant (dgemm order)
Mat small_strain_K(int dim, const Mat& dN, const Mat& D, bool bvar) {
    int n = (int)dN.size(), nd = n * dim, s = dim == 2 ? 3 : 6;
    Mat K = zeros(nd, nd);
    if (bvar) {
        Mat B = zeros(s, nd);
        for (int a = 0; a < n; ++a)
            for (int i = 0; i < dim; ++i) B[i][dim * a + i] = dN[a][i];
        for (int a = 0; a < n; ++a) {
            if (dim == 2) { B[2][2 * a] = dN[a][1]; B[2][2 * a + 1] = dN[a][0]; }
            else {
                B[3][3 * a + 1] = dN[a][2]; B[3][3 * a + 2] = dN[a][1];
                B[4][3 * a] = dN[a][2]; B[4][3 * a + 2] = dN[a][0];
                B[5][3 * a] = dN[a][1]; B[5][3 * a + 1] = dN[a][0];
            }
        }
        Mat BtD = zeros(nd, s);
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < s; ++j)
                for (int k = 0; k < s; ++k) BtD[i][j] += B[k][i] * D[k][j];
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j)
                for (int k = 0; k < s; ++k) K[i][j] += BtD[i][k] * B[k][j];
        return K;
    }
    for (int row = 0; row < nd; ++row) {
        int is = row % dim, in = row / dim;
        double pt[6];
        const auto& g = dN[in];
        if (dim == 2) {
            for (int k = 0; k < 3; ++k) pt[k] = is == 0 ? g[0] * D[0][k] + g[1] * D[2][k] : g[1] * D[1][k] + g[0] * D[2][k];
        } else {
            for (int k = 0; k < 6; ++k)
                pt[k] = is == 0 ? g[0] * D[0][k] + g[2] * D[4][k] + g[1] * D[5][k]
                      : is == 1 ? g[1] * D[1][k] + g[2] * D[3][k] + g[0] * D[5][k]
                                : g[2] * D[2][k] + g[1] * D[3][k] + g[0] * D[4][k];
        }
        for (int col = 0; col < nd; ++col) {
            int js = col % dim;
            const auto& h = dN[col / dim];
            if (dim == 2) K[row][col] = js == 0 ? pt[0] * h[0] + pt[2] * h[1] : pt[1] * h[1] + pt[2] * h[0];
            else K[row][col] = js == 0 ? pt[0] * h[0] + pt[4] * h[2] + pt[5] * h[1]
                             : js == 1 ? pt[1] * h[1] + pt[3] * h[2] + pt[5] * h[0]
                                       : pt[2] * h[2] + pt[3] * h[1] + pt[4] * h[0];
        }
    }
    return K;
}

// threads of the element loop (orc_set_threads); 1 = the serial reference loop
int g_threads = 1;

struct Assembled {
    std::vector<std::map<int, double>> rows;  // all dof rows
    Vec P;
};

// element K_e, P_e (assemble_elem_nonlinear), writes trial delta_max and flags
// absm: instead of K_e, the bound sum |w| |B|^T |D| |B| (and |w| |N| |N| |R|^T |D_c| |R|
// for cohesive points) of the absolute values of every product that enters an
// entry; the rounding error of any summation order of K_e is below a small
// multiple of the unit roundoff times this sum.
void element(const Problem& Pb, int64_t e, const Vec& U, double dt, const Vec& dconv, Vec& dtrial,
             std::vector<int>& flags, bool bvar, Mat& Ke, Vec& Pe, bool absm = false) {
    const int dim = Pb.dim;
    const auto& nodes = Pb.elem_nodes[e];
    const int nt = (int)nodes.size(), m = nt * dim, s = dim == 2 ? 3 : 6;
    Vec ue(m);
    for (int n = 0; n < nt; ++n)
        for (int d = 0; d < dim; ++d) ue[n * dim + d] = U[(size_t)d * Pb.n_all + nodes[n]];
    Ke = zeros(m, m);
    Pe.assign(m, 0.0);
    for (const QP& q : Pb.qps[e]) {
        if (q.coh) continue;
        const Mat& D = Pb.D[q.mat];
        Mat Kq;
        if (absm) {
            Mat aD = D, aG = q.grad;
            for (auto& r : aD) for (double& v : r) v = std::fabs(v);
            for (auto& r : aG) for (double& v : r) v = std::fabs(v);
            Kq = small_strain_K(dim, aG, aD, false);
        } else {
            Kq = small_strain_K(dim, q.grad, D, bvar);
        }
        const double wq = absm ? std::fabs(q.w) : q.w;
        Vec eps(s, 0.0), sig(s, 0.0);
        for (int n = 0; n < nt; ++n) {
            const auto& g = q.grad[n];
            if (dim == 2) {
                eps[0] += g[0] * ue[n * 2]; eps[1] += g[1] * ue[n * 2 + 1]; eps[2] += g[1] * ue[n * 2] + g[0] * ue[n * 2 + 1];
            } else {
                eps[0] += g[0] * ue[n * 3]; eps[1] += g[1] * ue[n * 3 + 1]; eps[2] += g[2] * ue[n * 3 + 2];
                eps[3] += g[2] * ue[n * 3 + 1] + g[1] * ue[n * 3 + 2];
                eps[4] += g[2] * ue[n * 3] + g[0] * ue[n * 3 + 2];
                eps[5] += g[1] * ue[n * 3] + g[0] * ue[n * 3 + 1];
            }
        }
        if (dim == 2) {
            sig[0] = D[0][0] * eps[0] + D[0][1] * eps[1]; sig[1] = D[1][0] * eps[0] + D[1][1] * eps[1]; sig[2] = D[2][2] * eps[2];
        } else {
            for (int i = 0; i < 3; ++i) sig[i] = D[i][0] * eps[0] + D[i][1] * eps[1] + D[i][2] * eps[2];
            for (int i = 3; i < 6; ++i) sig[i] = D[i][i] * eps[i];
        }
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) Ke[i][j] += wq * Kq[i][j];
        for (int n = 0; n < nt; ++n) {
            const auto& g = q.grad[n];
            double f[3];
            if (dim == 2) { f[0] = g[0] * sig[0] + g[1] * sig[2]; f[1] = g[1] * sig[1] + g[0] * sig[2]; }
            else {
                f[0] = g[0] * sig[0] + g[2] * sig[4] + g[1] * sig[5];
                f[1] = g[1] * sig[1] + g[2] * sig[3] + g[0] * sig[5];
                f[2] = g[2] * sig[2] + g[1] * sig[3] + g[0] * sig[4];
            }
            for (int d = 0; d < dim; ++d) Pe[n * dim + d] += q.w * f[d];
        }
    }
    // cohesive points grouped per cohesive element: K_coh accumulated then added
    size_t i0 = 0;
    const auto& Q = Pb.qps[e];
    while (i0 < Q.size()) {
        if (!Q[i0].coh) { ++i0; continue; }
        size_t i1 = i0;
        while (i1 < Q.size() && Q[i1].coh && Q[i1].map == Q[i0].map) ++i1;
        const auto& map = Q[i0].map;
        const int nh = (int)map.size() / 2, mc = 2 * nh * dim;
        Mat Kc = zeros(mc, mc);
        Vec Pc(mc, 0.0);
        Vec uc(mc);
        for (int n = 0; n < 2 * nh; ++n)
            for (int d = 0; d < dim; ++d) uc[n * dim + d] = ue[map[n] * dim + d];
        for (size_t k = i0; k < i1; ++k) {
            const QP& q = Q[k];
            Vec dl(dim, 0.0), delta(dim, 0.0);
            for (int d = 0; d < dim; ++d)
                for (int n = 0; n < nh; ++n) dl[d] += q.N[n] * (uc[(n + nh) * dim + d] - uc[n * dim + d]);
            for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j) delta[i] += q.R[i][j] * dl[j];
            Vec trac;
            Mat Dc;
            double dn;
            int fl;
            law_eval(Pb.laws[q.mat], dim, delta, dconv[q.isv], dt, trac, Dc, dn, fl);
            dtrial[q.isv] = dn;
            flags[q.isv] = fl;
            Mat RtD = zeros(dim, dim), RDR = zeros(dim, dim);
            auto ab = [&](double v) { return absm ? std::fabs(v) : v; };
            for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j)
                    for (int kk = 0; kk < dim; ++kk) RtD[i][j] += ab(q.R[kk][i]) * ab(Dc[kk][j]);
            for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j)
                    for (int kk = 0; kk < dim; ++kk) RDR[i][j] += RtD[i][kk] * ab(q.R[kk][j]);
            Vec tR(dim, 0.0);
            for (int i = 0; i < dim; ++i)
                for (int kk = 0; kk < dim; ++kk) tR[i] += q.R[kk][i] * trac[kk];
            for (int n1 = 0; n1 < nh; ++n1)
                for (int n2 = 0; n2 < nh; ++n2)
                    for (int d1 = 0; d1 < dim; ++d1)
                        for (int d2 = 0; d2 < dim; ++d2) {
                            double v = q.N[n1] * q.N[n2] * RDR[d1][d2];
                            if (absm) {
                                const double a = std::fabs(q.w * v);
                                Kc[n1 * dim + d1][n2 * dim + d2] += a;
                                Kc[(n1 + nh) * dim + d1][n2 * dim + d2] += a;
                                Kc[n1 * dim + d1][(n2 + nh) * dim + d2] += a;
                                Kc[(n1 + nh) * dim + d1][(n2 + nh) * dim + d2] += a;
                                continue;
                            }
                            Kc[n1 * dim + d1][n2 * dim + d2] += q.w * v;
                            Kc[(n1 + nh) * dim + d1][n2 * dim + d2] += q.w * -v;
                            Kc[n1 * dim + d1][(n2 + nh) * dim + d2] += q.w * -v;
                            Kc[(n1 + nh) * dim + d1][(n2 + nh) * dim + d2] += q.w * v;
                        }
            for (int n = 0; n < nh; ++n)
                for (int d = 0; d < dim; ++d) {
                    double v = q.N[n] * tR[d];
                    Pc[n * dim + d] += q.w * -v;
                    Pc[(n + nh) * dim + d] += q.w * v;
                }
        }
        for (int a = 0; a < 2 * nh; ++a)
            for (int d1 = 0; d1 < dim; ++d1) {
                int r = map[a] * dim + d1;
                for (int b = 0; b < 2 * nh; ++b)
                    for (int d2 = 0; d2 < dim; ++d2) Ke[r][map[b] * dim + d2] += Kc[a * dim + d1][b * dim + d2];
                Pe[r] += Pc[a * dim + d1];
            }
        i0 = i1;
    }
}

// Periodic node sets of a structured grid as Mesh::build_periodic_nodesets
// (src/Mesh.cpp:603-741) lists them: corners, then (2D) bottom-top and
// left-right pairs; (3D) x/y/z edge quadruples, then bottom-top, left-right and
// front-back face pairs.
std::vector<std::vector<int>> periodic_sets(int dim, int Nx, int Ny, int Nz) {
    std::vector<std::vector<int>> S;
    if (dim == 2) {
        S.push_back({0, Nx, (Nx + 1) * Ny, (Nx + 1) * (Ny + 1) - 1});
        for (int n = 1; n < Nx; ++n) S.push_back({n, (Nx + 1) * Ny + n});
        for (int n = 1; n < Ny; ++n) S.push_back({n * (Nx + 1), (n + 1) * (Nx + 1) - 1});
        return S;
    }
    const int top = (Nx + 1) * (Ny + 1) * Nz, lay = (Nx + 1) * (Ny + 1);
    S.push_back({0, Nx, (Nx + 1) * Ny, lay - 1, top, top + Nx, top + (Nx + 1) * Ny, top + lay - 1});
    for (int n = 1; n < Nx; ++n) S.push_back({n, (Nx + 1) * Ny + n, n + top, (Nx + 1) * Ny + n + top});
    for (int n = 1; n < Ny; ++n) S.push_back({n * (Nx + 1), (n + 1) * (Nx + 1) - 1, n * (Nx + 1) + top, (n + 1) * (Nx + 1) - 1 + top});
    for (int n = 1; n < Nz; ++n) S.push_back({n * lay, n * lay + Nx, n * lay + (Nx + 1) * Ny, n * lay + lay - 1});
    for (int a = 1; a < Nx; ++a)
        for (int b = 1; b < Ny; ++b) S.push_back({b * (Nx + 1) + a, b * (Nx + 1) + a + top});
    for (int a = 1; a < Ny; ++a)
        for (int b = 1; b < Nz; ++b) S.push_back({a * (Nx + 1) + b * lay, a * (Nx + 1) + b * lay + Nx});
    for (int a = 1; a < Nx; ++a)
        for (int b = 1; b < Nz; ++b) S.push_back({a + b * lay, a + b * lay + Ny * (Nx + 1)});
    return S;
}

// DofObject::distribute_dofs: free dofs in node order, prescribed after all
// free ones; with periodic sets only the primary node (set[0]) is numbered and
// only its Dirichlet conditions count, the other members copy its dofs
// (set_local_owned_normal_dofs / set_periodic_nodes, src/DofObject.cpp:336-495)
void number_dofs(Problem& P, const int32_t* bn, const int32_t* bd, const double* bv, int nbc) {
    std::vector<char> isp((size_t)P.n_all * P.dim, 0);
    Vec val((size_t)P.n_all * P.dim, 0.0);
    for (int i = 0; i < nbc; ++i) { isp[(size_t)bn[i] * P.dim + bd[i]] = 1; val[(size_t)bn[i] * P.dim + bd[i]] = bv[i]; }
    std::vector<char> secondary(P.n_all, 0);
    for (const auto& set : P.psets)
        for (size_t k = 1; k < set.size(); ++k) secondary[set[k]] = 1;
    P.n_free = 0;
    P.n_pres = 0;
    for (int64_t n = 0; n < P.n_all; ++n)
        if (!secondary[n])
            for (int d = 0; d < P.dim; ++d) (isp[(size_t)n * P.dim + d] ? P.n_pres : P.n_free)++;
    P.dofs.assign(P.n_all, std::vector<int>(P.dim));
    int64_t f = 0, p = 0;
    P.pres_val.assign(P.n_pres, 0.0);
    for (int64_t n = 0; n < P.n_all; ++n) {
        if (secondary[n]) continue;
        for (int d = 0; d < P.dim; ++d) {
            if (isp[(size_t)n * P.dim + d]) { P.pres_val[p] = val[(size_t)n * P.dim + d]; P.dofs[n][d] = (int)(P.n_free + p++); }
            else P.dofs[n][d] = (int)f++;
        }
    }
    for (const auto& set : P.psets)
        for (size_t k = 1; k < set.size(); ++k) P.dofs[set[k]] = P.dofs[set[0]];
}

Assembled assemble(const Problem& Pb, const Vec& U, double dt, const Vec& dconv, Vec& dtrial, std::vector<int>& flags,
                   bool bvar, bool absm = false) {
    Assembled A;
    A.rows.resize(Pb.n_free + Pb.n_pres);
    A.P.assign(Pb.n_free + Pb.n_pres, 0.0);
    dtrial = dconv;
    flags.assign(Pb.n_isv, 0);
    auto scatter = [&](int64_t e, const Mat& Ke, const Vec& Pe) {
        std::vector<int> ed;
        for (int n : Pb.elem_nodes[e])
            for (int d = 0; d < Pb.dim; ++d) ed.push_back(Pb.dofs[n][d]);
        for (size_t i = 0; i < ed.size(); ++i) {
            for (size_t j = 0; j < ed.size(); ++j) A.rows[ed[i]][ed[j]] += absm ? std::fabs(Ke[i][j]) : Ke[i][j];
            if (Pe[i] != 0.0) A.P[ed[i]] += Pe[i];
        }
    };
    if (g_threads <= 1) {
        for (int64_t e = Pb.e_begin; e < Pb.e_end; ++e) {
            Mat Ke;
            Vec Pe;
            element(Pb, e, U, dt, dconv, dtrial, flags, bvar, Ke, Pe, absm);
            scatter(e, Ke, Pe);
        }
        return A;
    }
    // element matrices of a chunk in parallel, scattered in element order by
    // one thread (the same sums as the serial loop)
    const int64_t chunk = 8192;
    std::vector<Mat> Kc(chunk);
    std::vector<Vec> Pc(chunk);
    for (int64_t e0 = Pb.e_begin; e0 < Pb.e_end; e0 += chunk) {
        const int64_t e1 = std::min(Pb.e_end, e0 + chunk);
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads)
        for (int64_t e = e0; e < e1; ++e) element(Pb, e, U, dt, dconv, dtrial, flags, bvar, Kc[e - e0], Pc[e - e0], absm);
        for (int64_t e = e0; e < e1; ++e) scatter(e, Kc[e - e0], Pc[e - e0]);
    }
    return A;
}

// Jacobi-PCG on K_ff (rows < n_free, cols < n_free)
int pcg(const std::vector<std::map<int, double>>& rows, int64_t nf, const Vec& b, Vec& x, double rtol, int maxit) {
    std::vector<int64_t> rp(nf + 1, 0);
    std::vector<int> ci;
    Vec cv, dinv(nf, 1.0);
    for (int64_t r = 0; r < nf; ++r) {
        for (auto& kv : rows[r])
            if (kv.first < nf) {
                ci.push_back(kv.first);
                cv.push_back(kv.second);
                if (kv.first == r && kv.second != 0) dinv[r] = 1.0 / kv.second;
            }
        rp[r + 1] = (int64_t)ci.size();
    }
    auto mv = [&](const Vec& v, Vec& y) {
        for (int64_t r = 0; r < nf; ++r) {
            double s = 0;
            for (int64_t j = rp[r]; j < rp[r + 1]; ++j) s += cv[j] * v[ci[j]];
            y[r] = s;
        }
    };
    x.assign(nf, 0.0);
    Vec r = b, z(nf), p(nf), q(nf);
    double bb = 0;
    for (double v : b) bb += v * v;
    if (bb == 0) return 0;
    for (int64_t i = 0; i < nf; ++i) { z[i] = dinv[i] * r[i]; p[i] = z[i]; }
    double rz = 0;
    for (int64_t i = 0; i < nf; ++i) rz += r[i] * z[i];
    int it = 0;
    for (; it < maxit; ++it) {
        mv(p, q);
        double pq = 0;
        for (int64_t i = 0; i < nf; ++i) pq += p[i] * q[i];
        double a = rz / pq, rr = 0, rzn = 0;
        for (int64_t i = 0; i < nf; ++i) {
            x[i] += a * p[i];
            r[i] -= a * q[i];
            z[i] = dinv[i] * r[i];
            rr += r[i] * r[i];
            rzn += r[i] * z[i];
        }
        if (rr <= rtol * rtol * bb) { ++it; break; }
        double beta = rzn / rz;
        rz = rzn;
        for (int64_t i = 0; i < nf; ++i) p[i] = z[i] + beta * p[i];
    }
    return it;
}

struct OrcHandle {
    Problem P;
    Assembled last;
    Vec U;            // [dim][n_all]
    Vec dconv, dtrial;
    std::vector<int> flags;
    Vec fext;         // [n_free] external load of the Newton residual (zero by default)
};

// Voigt strain and stress of a volumetric point (SmallStrainFast and
// computeStress for the isotropic D)
void point_strain(int dim, const QP& q, const Vec& ue, Vec& eps) {
    const int nt = (int)q.grad.size();
    eps.assign(dim == 2 ? 3 : 6, 0.0);
    for (int n = 0; n < nt; ++n) {
        const auto& g = q.grad[n];
        if (dim == 2) {
            eps[0] += g[0] * ue[n * 2]; eps[1] += g[1] * ue[n * 2 + 1]; eps[2] += g[1] * ue[n * 2] + g[0] * ue[n * 2 + 1];
        } else {
            eps[0] += g[0] * ue[n * 3]; eps[1] += g[1] * ue[n * 3 + 1]; eps[2] += g[2] * ue[n * 3 + 2];
            eps[3] += g[2] * ue[n * 3 + 1] + g[1] * ue[n * 3 + 2];
            eps[4] += g[2] * ue[n * 3] + g[0] * ue[n * 3 + 2];
            eps[5] += g[1] * ue[n * 3] + g[0] * ue[n * 3 + 1];
        }
    }
}
Vec mat_vec(const Mat& D, const Vec& v) {
    Vec o(D.size(), 0.0);
    for (size_t i = 0; i < D.size(); ++i)
        for (size_t j = 0; j < v.size(); ++j) o[i] += D[i][j] * v[j];
    return o;
}
// B^T s for every node of the point (SmallStrainInternalForceFast), into f[nt*dim]
void point_BtS(int dim, const QP& q, const Vec& s, double w, Vec& f) {
    const int nt = (int)q.grad.size();
    for (int n = 0; n < nt; ++n) {
        const auto& g = q.grad[n];
        if (dim == 2) {
            f[n * 2] += w * (g[0] * s[0] + g[1] * s[2]);
            f[n * 2 + 1] += w * (g[1] * s[1] + g[0] * s[2]);
        } else {
            f[n * 3] += w * (g[0] * s[0] + g[2] * s[4] + g[1] * s[5]);
            f[n * 3 + 1] += w * (g[1] * s[1] + g[2] * s[3] + g[0] * s[5]);
            f[n * 3 + 2] += w * (g[2] * s[2] + g[1] * s[3] + g[0] * s[4]);
        }
    }
}

}  // namespace

extern "C" {

static int64_t g_e_begin = -1, g_e_end = -1;
static int g_periodic[3] = {0, 0, 0};

// The next orc_create treats the mesh as the periodic RVE of an Nx x Ny (x Nz)
// generated grid; 0, 0, 0 switches periodicity off.
ORC_API void orc_periodic_grid(int Nx, int Ny, int Nz) {
    g_periodic[0] = Nx;
    g_periodic[1] = Ny;
    g_periodic[2] = Nz;
}

// Restrict the next orc_create to elements [e0, e1) (tables and assembly):
// used only to time a bounded sample of a large mesh.  -1, -1 restores the
// whole mesh.
ORC_API void orc_sample_range(int64_t e0, int64_t e1) {
    g_e_begin = e0;
    g_e_end = e1;
}

// cut: n_cut records of 128 int32 (pgfh_get_cut layout); laws: [n_law][7] = type, sc, dc, df, beta, visc, tfloor
ORC_API void* orc_create(int dim, int plane_strain, double tet_scale, int64_t n_nodes, int64_t n_all,
                         const double* coords, int64_t n_elem, const int32_t* conn, const int32_t* elem_mat, int n_mat,
                         const double* E, const double* nu, int n_law, const double* laws, int64_t n_cut,
                         const int32_t* cut, int n_bc, const int32_t* bc_node, const int32_t* bc_dof,
                         const double* bc_val) {
    auto* h = new OrcHandle();
    Problem& P = h->P;
    P.dim = dim;
    P.plane_strain = plane_strain;
    P.tet_scale = tet_scale;
    P.n_nodes = n_nodes;
    P.n_all = n_all;
    P.n_elem = n_elem;
    P.X.assign(n_all, Vec(3));
    for (int64_t n = 0; n < n_all; ++n)
        for (int d = 0; d < 3; ++d) P.X[n][d] = coords[d * n_all + n];
    P.elem_nodes.resize(n_elem);
    P.elem_mat.assign(elem_mat, elem_mat + n_elem);
    for (int m = 0; m < n_mat; ++m) P.D.push_back(isotropic_D(dim, plane_strain != 0, E[m], nu[m]));
    for (int k = 0; k < n_law; ++k) {
        const double* l = laws + 7 * k;
        P.laws.push_back(Law{(int)l[0], l[1], l[2], l[3], l[4], l[5], l[6]});
    }
    for (int64_t e = 0; e < n_elem; ++e) P.elem_nodes[e] = {conn[e * 4], conn[e * 4 + 1], conn[e * 4 + 2], conn[e * 4 + 3]};
    P.e_begin = g_e_begin < 0 ? 0 : g_e_begin;
    P.e_end = g_e_end < 0 ? n_elem : std::min<int64_t>(g_e_end, n_elem);
    if (g_periodic[0] > 0) P.psets = periodic_sets(dim, g_periodic[0], g_periodic[1], g_periodic[2]);
    build_tables(P, cut, n_cut);
    number_dofs(P, bc_node, bc_dof, bc_val, n_bc);
    h->U.assign((size_t)dim * n_all, 0.0);
    h->dconv.assign(P.n_isv, 0.0);
    h->dtrial = h->dconv;
    return h;
}

ORC_API void orc_destroy(void* h) { delete (OrcHandle*)h; }

// threads used by the element loop of orc_assemble (CPU-baseline timing)
ORC_API void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

// Jacobi-PCG iterations on a caller-supplied CSR K_ff (rows [0, n), the first
// rowsplit[r] entries of each row are its free columns): runs exactly `iters`
// iterations from x = 0 for the right-hand side b with `threads` OpenMP
// threads and returns the wall time per iteration in ms (CPU-baseline timing
// of the Krylov loop; the arithmetic is the oracle's pcg()).
ORC_API double orc_pcg_csr_ms_per_iter(int64_t n, const int64_t* rowptr, const int32_t* rowsplit, const int32_t* col,
                                       const double* val, const double* b, int iters, int threads) {
    if (threads < 1) threads = 1;
    Vec x(n, 0.0), r(b, b + n), z(n), p(n), q(n), dinv(n, 1.0);
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = rowptr[i]; j < rowptr[i] + rowsplit[i]; ++j)
            if (col[j] == i && val[j] != 0.0) dinv[i] = 1.0 / val[j];
    double rz = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : rz)
    for (int64_t i = 0; i < n; ++i) {
        z[i] = dinv[i] * r[i];
        p[i] = z[i];
        rz += r[i] * z[i];
    }
    auto t0 = std::chrono::steady_clock::now();
    for (int it = 0; it < iters; ++it) {
        double pq = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : pq)
        for (int64_t i = 0; i < n; ++i) {
            double s = 0;
            for (int64_t j = rowptr[i]; j < rowptr[i] + rowsplit[i]; ++j) s += val[j] * p[col[j]];
            q[i] = s;
            pq += p[i] * s;
        }
        const double a = rz / pq;
        double rzn = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : rzn)
        for (int64_t i = 0; i < n; ++i) {
            x[i] += a * p[i];
            r[i] -= a * q[i];
            z[i] = dinv[i] * r[i];
            rzn += r[i] * z[i];
        }
        const double beta = rzn / rz;
        rz = rzn;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / iters;
}

ORC_API void orc_sizes(void* hh, int64_t* out) {
    auto* h = (OrcHandle*)hh;
    out[0] = h->P.n_free;
    out[1] = h->P.n_pres;
    out[2] = h->P.n_isv;
    int64_t nq = 0;
    for (auto& q : h->P.qps) nq += (int64_t)q.size();
    out[3] = nq;
}

ORC_API void orc_dofs(void* hh, int32_t* dof, double* pres) {
    auto* h = (OrcHandle*)hh;
    for (int64_t n = 0; n < h->P.n_all; ++n)
        for (int d = 0; d < h->P.dim; ++d) dof[d * h->P.n_all + n] = h->P.dofs[n][d];
    for (int64_t k = 0; k < h->P.n_pres; ++k) pres[k] = h->P.pres_val[k];
}

// Sparsity from populate_local_neighbors: per node the set of dofs of every
// element sharing it; returned as full CSR over all rows (sorted columns).
ORC_API int64_t orc_pattern(void* hh, int64_t* rowptr, int32_t* colidx) {
    auto* h = (OrcHandle*)hh;
    const Problem& P = h->P;
    std::vector<std::set<int>> nb(P.n_all);
    for (int64_t e = 0; e < P.n_elem; ++e)
        for (int a : P.elem_nodes[e])
            for (int b : P.elem_nodes[e])
                for (int d = 0; d < P.dim; ++d) nb[a].insert(P.dofs[b][d]);
    // Solver::combine_periodic_contributions (src/Solver.cpp:832-847)
    for (const auto& set : P.psets)
        for (size_t k = 1; k < set.size(); ++k) nb[set[0]].insert(nb[set[k]].begin(), nb[set[k]].end());
    const int64_t n = P.n_free + P.n_pres;
    std::vector<int> rnode(n, -1);
    for (int64_t a = 0; a < P.n_all; ++a)
        for (int d = 0; d < P.dim; ++d)
            if (rnode[P.dofs[a][d]] < 0) rnode[P.dofs[a][d]] = (int)a;
    int64_t nnz = 0;
    if (rowptr) rowptr[0] = 0;
    for (int64_t r = 0; r < n; ++r) {
        for (int c : nb[rnode[r]]) {
            if (colidx) colidx[nnz] = c;
            ++nnz;
        }
        if (rowptr) rowptr[r + 1] = nnz;
    }
    return nnz;
}

ORC_API void orc_set_solution(void* hh, const double* U) {
    auto* h = (OrcHandle*)hh;
    std::memcpy(h->U.data(), U, sizeof(double) * h->U.size());
}
ORC_API void orc_get_solution(void* hh, double* U) {
    auto* h = (OrcHandle*)hh;
    std::memcpy(U, h->U.data(), sizeof(double) * h->U.size());
}

// assemble at the stored solution; values returned in the CSR order of orc_pattern
ORC_API double orc_assemble(void* hh, double dt, int bvariant, double* vals, double* P) {
    auto* h = (OrcHandle*)hh;
    auto t0 = std::chrono::steady_clock::now();
    h->last = assemble(h->P, h->U, dt, h->dconv, h->dtrial, h->flags, bvariant != 0);
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (vals) {
        int64_t k = 0;
        for (auto& row : h->last.rows)
            for (auto& kv : row) vals[k++] = kv.second;
    }
    if (P) std::memcpy(P, h->last.P.data(), sizeof(double) * h->last.P.size());
    return ms;
}

// the absolute-value bound of every K entry (see element()), in the CSR order of
// orc_pattern, at the stored solution; the trial state is left unchanged
ORC_API void orc_assemble_abs(void* hh, double dt, double* vals) {
    auto* h = (OrcHandle*)hh;
    Vec dtrial = h->dtrial;
    std::vector<int> flags = h->flags;
    Assembled A = assemble(h->P, h->U, dt, h->dconv, dtrial, flags, false, true);
    int64_t k = 0;
    for (auto& row : A.rows)
        for (auto& kv : row) vals[k++] = kv.second;
}

ORC_API void orc_isv(void* hh, double* dtrial, double* dconv, int32_t* flags) {
    auto* h = (OrcHandle*)hh;
    for (int64_t i = 0; i < h->P.n_isv; ++i) {
        if (dtrial) dtrial[i] = h->dtrial[i];
        if (dconv) dconv[i] = h->dconv[i];
        if (flags) flags[i] = h->flags.empty() ? 0 : h->flags[i];
    }
}
ORC_API void orc_isv_commit(void* hh) {
    auto* h = (OrcHandle*)hh;
    h->dconv = h->dtrial;
}

ORC_API int orc_pcg_last(void* hh, const double* b, double* x, double rtol, int maxit) {
    auto* h = (OrcHandle*)hh;
    Vec bv(b, b + h->P.n_free), xv;
    int it = pcg(h->last.rows, h->P.n_free, bv, xv, rtol, maxit);
    std::memcpy(x, xv.data(), sizeof(double) * h->P.n_free);
    return it;
}

// Newton load step (SolverNonlinearKSP::nonlinearStep) at load fraction frac.
// Returns Newton iterations (negative if not converged); info = res0, res, ms.
ORC_API int orc_newton_step(void* hh, double frac, double dt, double rel_tol, double abs_tol, int max_iter,
                            double lin_rtol, int bvariant, double* info) {
    auto* h = (OrcHandle*)hh;
    Problem& P = h->P;
    auto t0 = std::chrono::steady_clock::now();
    for (int64_t n = 0; n < P.n_all; ++n)
        for (int d = 0; d < P.dim; ++d) {
            int g = P.dofs[n][d];
            if (g >= P.n_free) h->U[(size_t)d * P.n_all + n] = frac * P.pres_val[g - P.n_free];
        }
    auto resid = [&](Vec& r) {
        h->last = assemble(P, h->U, dt, h->dconv, h->dtrial, h->flags, bvariant != 0);
        r.assign(P.n_free, 0.0);
        double s = 0;
        for (int64_t i = 0; i < P.n_free; ++i) {
            r[i] = (h->fext.empty() ? 0.0 : frac * h->fext[i]) - h->last.P[i];
            s += r[i] * r[i];
        }
        return std::sqrt(s);
    };
    Vec r;
    double res0 = resid(r), res = res0;
    int it = 0;
    bool conv = res0 < abs_tol || res0 == 0.0;
    while (!conv && it < max_iter) {
        Vec dx;
        pcg(h->last.rows, P.n_free, r, dx, lin_rtol, 200000);
        for (int64_t n = 0; n < P.n_all; ++n)
            for (int d = 0; d < P.dim; ++d) {
                int g = P.dofs[n][d];
                if (g < P.n_free) h->U[(size_t)d * P.n_all + n] += dx[g];
            }
        ++it;
        res = resid(r);
        if (res < rel_tol * res0 || res < abs_tol) conv = true;
        if (!std::isfinite(res)) break;
    }
    if (info) {
        info[0] = res0;
        info[1] = res;
        info[2] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return conv ? it : -it - 1;
}

// external load of the Newton residual r = frac * fext - P_int (free rows)
ORC_API void orc_set_fext(void* hh, const double* fext) {
    auto* h = (OrcHandle*)hh;
    h->fext.assign(fext, fext + h->P.n_free);
}

// Macro-strain body load of a linear RVE, BodyLoadMacroStrainLinear::BF_Kernel
// (src/BodyLoadMacroStrainLinear.cpp:10-30): per volumetric point
// -W J B^T (D eps_bar), assembled over all rows (Assembler::set_vec order)
ORC_API void orc_macro_strain_load(void* hh, const double* eps_bar, double* f_all) {
    auto* h = (OrcHandle*)hh;
    const Problem& P = h->P;
    const int dim = P.dim, S = dim == 2 ? 3 : 6;
    Vec eb(eps_bar, eps_bar + S);
    std::fill(f_all, f_all + P.n_free + P.n_pres, 0.0);
    for (int64_t e = P.e_begin; e < P.e_end; ++e) {
        const int nt = (int)P.elem_nodes[e].size();
        Vec fe((size_t)nt * dim, 0.0);
        for (const QP& q : P.qps[e]) {
            if (q.coh) continue;
            Vec r1 = mat_vec(P.D[q.mat], eb);
            for (double& v : r1) v *= -1.0 * q.w;
            point_BtS(dim, q, r1, 1.0, fe);
        }
        for (int n = 0; n < nt; ++n)
            for (int d = 0; d < dim; ++d) f_all[P.dofs[P.elem_nodes[e][n]][d]] += fe[n * dim + d];
    }
}

// SubscaleModel::homogenize_stress_strain (src/SubscaleModel.cpp:42-196) at the
// stored solution (the periodic fluctuation): volume averages of the stress
// D (eps(u) + eps_bar), of the fluctuation strain eps(u) and of the energy
// density sigma : (eps(u) + eps_bar) over every volumetric point, and the
// volume sum W J.
ORC_API void orc_homogenize(void* hh, const double* eps_bar, double* avg_stress, double* avg_strain, double* volume,
                            double* avg_energy) {
    auto* h = (OrcHandle*)hh;
    const Problem& P = h->P;
    const int dim = P.dim, S = dim == 2 ? 3 : 6;
    Vec sig_sum(S, 0.0), eps_sum(S, 0.0);
    double vol = 0.0, en = 0.0;
    for (int64_t e = P.e_begin; e < P.e_end; ++e) {
        const auto& nodes = P.elem_nodes[e];
        const int nt = (int)nodes.size();
        Vec ue((size_t)nt * dim);
        for (int n = 0; n < nt; ++n)
            for (int d = 0; d < dim; ++d) ue[n * dim + d] = h->U[(size_t)d * P.n_all + nodes[n]];
        for (const QP& q : P.qps[e]) {
            if (q.coh) continue;
            Vec eps;
            point_strain(dim, q, ue, eps);
            Vec tot = eps;
            for (int i = 0; i < S; ++i) tot[i] += eps_bar[i];
            Vec sig = mat_vec(P.D[q.mat], tot);
            for (int i = 0; i < S; ++i) {
                sig_sum[i] += q.w * sig[i];
                eps_sum[i] += q.w * eps[i];
                en += q.w * sig[i] * tot[i];
            }
            vol += q.w;
        }
    }
    for (int i = 0; i < S; ++i) {
        avg_stress[i] = sig_sum[i] / vol;
        avg_strain[i] = eps_sum[i] / vol;
    }
    *volume = vol;
    // volume average of sigma : eps (Voigt, engineering shear)
    if (avg_energy) *avg_energy = en / vol;
}

// single evaluation of the traction-separation law: law = [type, sc, dc, df,
// beta, visc, tfloor]; out = trac[dim], Dc[dim*dim], dmax_new, flag
ORC_API void orc_law(const double* law, int dim, const double* delta_in, double dmax, double dt, double* out) {
    Law L{(int)law[0], law[1], law[2], law[3], law[4], law[5], law[6]};
    Vec delta(delta_in, delta_in + dim), trac;
    Mat Dc;
    double dn;
    int fl;
    law_eval(L, dim, delta, dmax, dt, trac, Dc, dn, fl);
    for (int i = 0; i < dim; ++i) out[i] = trac[i];
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j) out[dim + i * dim + j] = Dc[i][j];
    out[dim + dim * dim] = dn;
    out[dim + dim * dim + 1] = fl;
}

// shape functions of a volumetric parent element at a parent point: dim 3 the
// linear tet, dim 2 with nn = 3 the triangle, nn = 4 the bilinear quad and
// nn >= 5 the Wachspress n-gon on its regular reference polygon.  Also returns
// the polygon's reference vertices (vtx [nn][2], nn >= 5 only) and, for
// nn >= 5, the points and weights of its fan rule (rule [4 nn][3]: xi, eta, w).
ORC_API void orc_shape(int dim, int nn, const double* xi, double* N, double* dN, double* vtx, double* rule) {
    Vec x(xi, xi + dim), Nv;
    Mat dNv;
    shape(dim, nn, x, Nv, dNv);
    for (int k = 0; k < nn; ++k) {
        N[k] = Nv[k];
        for (int i = 0; i < dim; ++i) dN[k * dim + i] = dNv[k][i];
    }
    if (dim == 2 && nn >= 5) {
        if (vtx)
            for (int k = 0; k < nn; ++k) poly_vertex(nn, k, vtx + 2 * k);
        if (rule) {
            Rule r = rule_poly(nn);
            for (size_t q = 0; q < r.w.size(); ++q) {
                rule[3 * q] = r.xi[q][0];
                rule[3 * q + 1] = r.xi[q][1];
                rule[3 * q + 2] = r.w[q];
            }
        }
    }
}

// --------- integer flag checks: node detection and enrichment-node set -------
// inclusions: [n][5] = type, cx, cy, cz, r.  coords [3][n_nodes] are the
// ORIGINAL (unmoved) coordinates; first_elem_node0/1 give the "other" node of
// the nudge rule.  Outputs detection per node and the moved coordinates.
ORC_API void orc_detect_nodes(int64_t n_nodes, double* coords, int n_incl, const double* incl,
                              const int32_t* other_node, int32_t* det) {
    for (int64_t n = 0; n < n_nodes; ++n) {
        int moved = 0;
        int res = -1;
        for (int j = n_incl - 1; j >= 0; --j) {
            const double* I = incl + 5 * j;
            double x = coords[n], y = coords[n_nodes + n], z = coords[2 * n_nodes + n];
            bool sph = I[0] != 0;
            double lo[3] = {I[1] - I[4], I[2] - I[4], sph ? I[3] - I[4] : -1.7976931348623157e308};
            double hi[3] = {I[1] + I[4], I[2] + I[4], sph ? I[3] + I[4] : 1.7976931348623157e308};
            if (!(lo[0] <= x && lo[1] <= y && lo[2] <= z && x <= hi[0] && y <= hi[1] && z <= hi[2])) continue;
            double val;
            if (!sph) val = (std::sqrt((x - I[1]) * (x - I[1]) + (y - I[2]) * (y - I[2])) - I[4]) / I[4];
            else val = std::pow((x - I[1]) / I[4], 2) + std::pow((y - I[2]) / I[4], 2) + std::pow((z - I[3]) / I[4], 2) - 1.0;
            if (val > 1e-12) continue;
            if (val < -1e-12) { res = j; break; }
            if (moved >= 5) { res = -2; break; }
            double nv[3] = {x - I[1], y - I[2], sph ? z - I[3] : 0.0};
            double mg = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
            int64_t o = other_node[n];
            double dx = x - coords[o], dy = y - coords[n_nodes + o], dz = z - coords[2 * n_nodes + o];
            double mv = 0.05 * std::sqrt(dx * dx + dy * dy + dz * dz);
            for (int d = 0; d < 3; ++d) coords[d * n_nodes + n] += mv * (nv[d] / mg);
            ++moved;
            j = n_incl;
        }
        det[n] = res;
    }
}

// closed-form intersection of the segment p0->p1 with the inclusion surface:
// root in [0,1] of |p0 + a (p1-p0) - c|^2 = r^2
ORC_API double orc_edge_root(const double* p0, const double* p1, const double* incl) {
    int nd = incl[0] != 0 ? 3 : 2;
    double A = 0, B = 0, Cc = -incl[4] * incl[4];
    for (int d = 0; d < nd; ++d) {
        double dd = p1[d] - p0[d], e = p0[d] - incl[1 + d];
        A += dd * dd;
        B += 2 * dd * e;
        Cc += e * e;
    }
    double disc = std::sqrt(std::max(0.0, B * B - 4 * A * Cc));
    // numerically stable pair of roots
    double qq = -0.5 * (B + (B >= 0 ? disc : -disc));
    double r1 = qq / A, r2 = Cc / qq;
    if (r1 >= -1e-12 && r1 <= 1 + 1e-12) return r1;
    return r2;
}

}  // extern "C"
