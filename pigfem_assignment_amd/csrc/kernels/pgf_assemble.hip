// Element assembly kernels (volumetric uncut, cut/IGFEM with cohesive points).
//
// Per quadrature point the math is the reference's ProblemUtilities fast
// kernels: strain B*u (SmallStrainFast, src/Utilities.cpp:300-361), stress
// D*eps with the isotropic zero structure (computeStress,
// src/material_lei.cpp:232-261), K += w * B^T D B in the zero-skipping
// SmallStrainKFast form (src/Utilities.cpp:196-296), P += w * B^T sigma
// (SmallStrainInternalForceFast, :364-408), cohesive opening / stiffness /
// force (cohesiveDeltaFast, cohesiveKFast, cohesiveInternalForceFast,
// :465-531).  The scatter is Assembler::set_matrix / set_vec
// (src/Assembler.cpp:811-879).
//
// Two launch schemes, both free of atomics:
//  * gather (default): one thread per node owns that node's CSR rows, zeroes
//    them and accumulates the contributions of its incident elements in
//    element order (the reference's summation order).  Uncut elements are
//    evaluated on the fly; cut elements are evaluated once per element by one
//    wavefront into a dense scratch block first.  Every K entry reaches HBM
//    about once; the read-modify-writes stay in L2.
//  * coloured: one thread per uncut element / one wavefront per cut element,
//    colour by colour, read-modify-write straight into the CSR rows.
#include "pgf_kernels.hpp"
#include "pgf_device.hpp"

// Multiply-adds are fused only within one expression (never across
// statements), so the element arithmetic, and with it every assembled bit,
// is the same in every kernel variant and gather layout that evaluates the
// same expressions.
#pragma clang fp contract(on)

namespace pgfk {

// An isotropic D (src/material_lei.cpp:52-97, every material of the path)
// has three distinct entries: a on the normal diagonal, b off the diagonal of
// the normal block, c on the shear diagonal (plane strain, plane stress and
// 3D alike).  The products below are the full D products with the zero terms
// left out; the nonzero terms keep the full products' order, so the results
// are the same while the registers hold three numbers instead of S*S.
struct Iso {
    double a, b, c;
};
template <int DIM>
__device__ inline Iso iso_of(const double* D) {
    constexpr int S = DIM == 2 ? 3 : 6;
    return Iso{D[0], D[1], D[S * S - 1]};
}

// parts[k] = sum_l B_l,(a,da) D_lk for Voigt row k, the reference's part0..5
template <int DIM>
__device__ inline void k_parts(const double* g, int da, const Iso& m, double* parts) {
    if (DIM == 2) {
        if (da == 0) { parts[0] = g[0] * m.a; parts[1] = g[0] * m.b; parts[2] = g[1] * m.c; }
        else         { parts[0] = g[1] * m.b; parts[1] = g[1] * m.a; parts[2] = g[0] * m.c; }
    } else {
        if (da == 0) {
            parts[0] = g[0] * m.a; parts[1] = g[0] * m.b; parts[2] = g[0] * m.b;
            parts[3] = 0.0;        parts[4] = g[2] * m.c; parts[5] = g[1] * m.c;
        } else if (da == 1) {
            parts[0] = g[1] * m.b; parts[1] = g[1] * m.a; parts[2] = g[1] * m.b;
            parts[3] = g[2] * m.c; parts[4] = 0.0;        parts[5] = g[0] * m.c;
        } else {
            parts[0] = g[2] * m.b; parts[1] = g[2] * m.b; parts[2] = g[2] * m.a;
            parts[3] = g[1] * m.c; parts[4] = g[0] * m.c; parts[5] = 0.0;
        }
    }
}
// The same parts for a run-time da inside a loop: the normal entry of the
// node's B column is g[da] (the others zero), the shear entries select
// without branching on da.
template <int DIM>
__device__ inline void k_parts_col(const double* g, int da, const Iso& m, double* parts) {
    const double gn = DIM == 2 ? (da == 0 ? g[0] : g[1]) : (da == 0 ? g[0] : (da == 1 ? g[1] : g[2]));
    if (DIM == 2) {
        parts[0] = gn * (da == 0 ? m.a : m.b);
        parts[1] = gn * (da == 1 ? m.a : m.b);
        parts[2] = (da == 0 ? g[1] : g[0]) * m.c;
    } else {
        parts[0] = gn * (da == 0 ? m.a : m.b);
        parts[1] = gn * (da == 1 ? m.a : m.b);
        parts[2] = gn * (da == 2 ? m.a : m.b);
        parts[3] = da == 0 ? 0.0 : (da == 1 ? g[2] : g[1]) * m.c;
        parts[4] = da == 1 ? 0.0 : (da == 0 ? g[2] : g[0]) * m.c;
        parts[5] = da == 2 ? 0.0 : (da == 0 ? g[1] : g[0]) * m.c;
    }
}
// K(row, (b,db)) from parts and column node gradient gb
template <int DIM>
__device__ inline double k_entry(const double* parts, const double* gb, int db) {
    if (DIM == 2) return db == 0 ? parts[0] * gb[0] + parts[2] * gb[1] : parts[1] * gb[1] + parts[2] * gb[0];
    if (db == 0) return parts[0] * gb[0] + parts[4] * gb[2] + parts[5] * gb[1];
    if (db == 1) return parts[1] * gb[1] + parts[3] * gb[2] + parts[5] * gb[0];
    return parts[2] * gb[2] + parts[3] * gb[1] + parts[4] * gb[0];
}
// strain = B u for nn nodes with gradients g[n*DIM+i] and element vector u[n*DIM+d]
template <int DIM>
__device__ inline void strain_of(int nn, const double* g, const double* u, double* eps) {
    if (DIM == 2) {
        eps[0] = eps[1] = eps[2] = 0.0;
        for (int n = 0; n < nn; ++n) eps[0] += g[n * 2 + 0] * u[n * 2];
        for (int n = 0; n < nn; ++n) eps[1] += g[n * 2 + 1] * u[n * 2 + 1];
        for (int n = 0; n < nn; ++n) eps[2] += g[n * 2 + 1] * u[n * 2] + g[n * 2 + 0] * u[n * 2 + 1];
    } else {
        for (int k = 0; k < 6; ++k) eps[k] = 0.0;
        for (int n = 0; n < nn; ++n) eps[0] += g[n * 3 + 0] * u[n * 3];
        for (int n = 0; n < nn; ++n) eps[1] += g[n * 3 + 1] * u[n * 3 + 1];
        for (int n = 0; n < nn; ++n) eps[2] += g[n * 3 + 2] * u[n * 3 + 2];
        for (int n = 0; n < nn; ++n) eps[3] += g[n * 3 + 2] * u[n * 3 + 1] + g[n * 3 + 1] * u[n * 3 + 2];
        for (int n = 0; n < nn; ++n) eps[4] += g[n * 3 + 2] * u[n * 3] + g[n * 3 + 0] * u[n * 3 + 2];
        for (int n = 0; n < nn; ++n) eps[5] += g[n * 3 + 1] * u[n * 3] + g[n * 3 + 0] * u[n * 3 + 1];
    }
}
template <int DIM>
__device__ inline void stress_of(const Iso& m, const double* e, double* s) {
    if (DIM == 2) {
        s[0] = m.a * e[0] + m.b * e[1];
        s[1] = m.b * e[0] + m.a * e[1];
        s[2] = m.c * e[2];
    } else {
        s[0] = m.a * e[0] + m.b * e[1] + m.b * e[2];
        s[1] = m.b * e[0] + m.a * e[1] + m.b * e[2];
        s[2] = m.b * e[0] + m.b * e[1] + m.a * e[2];
        s[3] = m.c * e[3];
        s[4] = m.c * e[4];
        s[5] = m.c * e[5];
    }
}
// B^T sigma for one node (SmallStrainInternalForceFast)
template <int DIM>
__device__ inline double force_entry(const double* g, int d, const double* s) {
    if (DIM == 2) return d == 0 ? g[0] * s[0] + g[1] * s[2] : g[1] * s[1] + g[0] * s[2];
    if (d == 0) return g[0] * s[0] + g[2] * s[4] + g[1] * s[5];
    if (d == 1) return g[1] * s[1] + g[2] * s[3] + g[0] * s[5];
    return g[2] * s[2] + g[1] * s[3] + g[0] * s[4];
}

// Geometry of an uncut parent element at its quadrature point q: QUAD4 2x2
// Gauss points (w = J*1) or the linear tet's single merged point (the
// five-point rule's weights -0.8 + 4*0.45 applied to the common B),
// Elem::ShapeFunctions at the point.
template <int DIM>
__device__ inline void qp_geometry(const double (*X)[DIM], double tet_scale, int q, double* g, double& w) {
    if (DIM == 2) {
        const double r0 = 0.57735026918962576450914878050196;
        const double r = (q % 2 == 0) ? -r0 : r0, s = (q / 2 == 0) ? -r0 : r0;
        const double dr[4][2] = {{-0.25 * (1.0 - s), -0.25 * (1.0 - r)},
                                 {0.25 * (1.0 - s), -0.25 * (1.0 + r)},
                                 {0.25 * (1.0 + s), 0.25 * (1.0 + r)},
                                 {-0.25 * (1.0 + s), 0.25 * (1.0 - r)}};
        double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
        for (int n = 0; n < 4; ++n) {
            J00 += dr[n][0] * X[n][0]; J01 += dr[n][0] * X[n][1];
            J10 += dr[n][1] * X[n][0]; J11 += dr[n][1] * X[n][1];
        }
        const double det = J00 * J11 - J01 * J10;
        const double rd = 1.0 / det;  // one division per point
        const double i00 = J11 * rd, i01 = -J01 * rd, i10 = -J10 * rd, i11 = J00 * rd;
        for (int n = 0; n < 4; ++n) {
            g[n * 2 + 0] = i00 * dr[n][0] + i01 * dr[n][1];
            g[n * 2 + 1] = i10 * dr[n][0] + i11 * dr[n][1];
        }
        w = det * 1.0;
    } else {
        double J[3][3];
        for (int i = 0; i < 3; ++i) {
            J[0][i] = X[1][i] - X[0][i];
            J[1][i] = X[2][i] - X[0][i];
            J[2][i] = X[3][i] - X[0][i];
        }
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        const double rd = 1.0 / det;  // one division per element
        double inv[3][3];
        inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * rd;
        inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * rd;
        inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * rd;
        inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * rd;
        inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * rd;
        inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * rd;
        inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * rd;
        inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * rd;
        inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * rd;
        const double dr[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int n = 0; n < 4; ++n)
            for (int i = 0; i < 3; ++i) g[n * 3 + i] = inv[i][0] * dr[n][0] + inv[i][1] * dr[n][1] + inv[i][2] * dr[n][2];
        double ws = det * -0.8;
        for (int k = 0; k < 4; ++k) ws += det * 0.45;
        w = ws * tet_scale;
    }
}
// Row (la, da) of an affine QUAD4 (a parallelogram: x0 - x1 + x2 - x3 = 0
// exactly, every element of a generated grid whose nodes were not nudged)
// integrated in closed form.  With J constant, the 2x2 Gauss sum of
// grad N_a (x) grad N_b is det J * inv M_ab inv^T with the reference moments
// M_ab = sum_q dN_a/dxi(q) dN_b/dxi(q)^T:
//   M00 = xi_a xi_b (1 + eta_a eta_b / 3) / 4,  M01 = xi_a eta_b / 4,
//   M10 = eta_a xi_b / 4,  M11 = eta_a eta_b (1 + xi_a xi_b / 3) / 4.
// The 2x2 rule integrates these products exactly (degree two per variable),
// so the entries equal the quadrature's up to rounding; the row's internal
// force is the row times the element displacements (linear elasticity).
__device__ inline bool quad_affine(const double (*X)[2]) {
    return (X[0][0] - X[1][0]) + (X[2][0] - X[3][0]) == 0.0 && (X[0][1] - X[1][1]) + (X[2][1] - X[3][1]) == 0.0;
}
template <int NR>
__device__ inline void quad_affine_rows(const double (*X)[2], int la, const int* da, const Iso& m, double (*Kr)[8]) {
    // constant Jacobian from the centre derivatives dN_n/dxi = (xi_n, eta_n) / 4
    const double J00 = 0.25 * ((X[1][0] - X[0][0]) + (X[2][0] - X[3][0]));
    const double J01 = 0.25 * ((X[1][1] - X[0][1]) + (X[2][1] - X[3][1]));
    const double J10 = 0.25 * ((X[2][0] - X[1][0]) + (X[3][0] - X[0][0]));
    const double J11 = 0.25 * ((X[2][1] - X[1][1]) + (X[3][1] - X[0][1]));
    const double det = J00 * J11 - J01 * J10;
    const double rd = 1.0 / det;
    const double i00 = J11 * rd, i01 = -J01 * rd, i10 = -J10 * rd, i11 = J00 * rd;
    const double xa = (la == 1 || la == 2) ? 1.0 : -1.0, ea = la >= 2 ? 1.0 : -1.0;
    const double third = 1.0 / 3.0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const double xb = (b == 1 || b == 2) ? 1.0 : -1.0, eb = b >= 2 ? 1.0 : -1.0;
        const double M00 = 0.25 * xa * xb * (1.0 + ea * eb * third), M01 = 0.25 * xa * eb;
        const double M10 = 0.25 * ea * xb, M11 = 0.25 * ea * eb * (1.0 + xa * xb * third);
        // T = M inv^T, I = det inv T: I_ij = integral of g_a,i g_b,j
        const double T00 = M00 * i00 + M01 * i01, T01 = M00 * i10 + M01 * i11;
        const double T10 = M10 * i00 + M11 * i01, T11 = M10 * i10 + M11 * i11;
        const double I00 = det * (i00 * T00 + i01 * T10), I01 = det * (i00 * T01 + i01 * T11);
        const double I10 = det * (i10 * T00 + i11 * T10), I11 = det * (i10 * T01 + i11 * T11);
        for (int h = 0; h < NR; ++h) {
            if (da[h] == 0) {
                Kr[h][b * 2 + 0] = m.a * I00 + m.c * I11;
                Kr[h][b * 2 + 1] = m.b * I01 + m.c * I10;
            } else {
                Kr[h][b * 2 + 0] = m.b * I10 + m.c * I01;
                Kr[h][b * 2 + 1] = m.a * I11 + m.c * I00;
            }
        }
    }
}

template <int DIM>
__device__ inline void uncut_geometry(const double (*X)[DIM], double tet_scale, double (*g)[4 * DIM], double* w) {
    constexpr int NQ = DIM == 2 ? 4 : 1;
    for (int q = 0; q < NQ; ++q) qp_geometry<DIM>(X, tet_scale, q, g[q], w[q]);
}

template <typename KT>
__device__ inline void kadd(const AsmArgs& a, int64_t pos, double v) {
    if (sizeof(KT) == 8) a.K[pos] += v;
    else a.K32[pos] += (float)v;
}

// ---------------------------------------------------------------------------
// Coloured scheme, uncut elements: one thread per element.
template <int DIM, bool DO_K, bool DO_F, typename KT>
__global__ __launch_bounds__(256) void k_uncut(AsmArgs a, int64_t begin, int64_t end) {
    const int64_t p = begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= end) return;
    constexpr int S = DIM == 2 ? 3 : 6;
    constexpr int NQ = DIM == 2 ? 4 : 1;
    int nd[4];
    double X[4][DIM];
    for (int k = 0; k < 4; ++k) {
        nd[k] = a.ue_conn[k * a.n_uncut + p];
        for (int i = 0; i < DIM; ++i) X[k][i] = a.X[i * a.n_all + nd[k]];
    }
    const int mat = a.ue_mat[p];
    const Iso D = iso_of<DIM>(a.Dmat + mat * 36);
    double g[NQ][4 * DIM], w[NQ];
    uncut_geometry<DIM>(X, a.tet_scale, g, w);
    int row_dof[4][DIM];
    for (int n = 0; n < 4; ++n)
        for (int d = 0; d < DIM; ++d) row_dof[n][d] = a.dof[d * a.n_all + nd[n]];
    if (DO_F) {
        double u[4 * DIM];
        for (int n = 0; n < 4; ++n)
            for (int d = 0; d < DIM; ++d) u[n * DIM + d] = a.U[d * a.n_all + nd[n]];
        double Pe[4 * DIM];
        for (int i = 0; i < 4 * DIM; ++i) Pe[i] = 0.0;
        for (int q = 0; q < NQ; ++q) {
            double eps[S], sig[S];
            strain_of<DIM>(4, g[q], u, eps);
            stress_of<DIM>(D, eps, sig);
            for (int n = 0; n < 4; ++n)
                for (int d = 0; d < DIM; ++d) Pe[n * DIM + d] += w[q] * force_entry<DIM>(&g[q][n * DIM], d, sig);
        }
        for (int n = 0; n < 4; ++n)
            for (int d = 0; d < DIM; ++d) {
                const int r = row_dof[n][d];
                if (r < a.n_free || a.all_rows) a.P[r] += Pe[n * DIM + d];
            }
    }
    if (DO_K) {
        uint32_t offw[8];
        for (int k = 0; k < 8; ++k) offw[k] = a.ue_off[k * a.n_uncut + p];
        for (int an = 0; an < 4; ++an)
            for (int da = 0; da < DIM; ++da) {
                const int r = row_dof[an][da];
                if (r >= a.n_free && !a.all_rows) continue;
                double Kr[4 * DIM];
                for (int j = 0; j < 4 * DIM; ++j) Kr[j] = 0.0;
                for (int q = 0; q < NQ; ++q) {
                    double parts[S];
                    k_parts<DIM>(&g[q][an * DIM], da, D, parts);
                    for (int bn = 0; bn < 4; ++bn)
                        for (int db = 0; db < DIM; ++db)
                            Kr[bn * DIM + db] += w[q] * k_entry<DIM>(parts, &g[q][bn * DIM], db);
                }
                const int64_t base = a.rowptr[r];
                for (int bn = 0; bn < 4; ++bn) {
                    const int k = (an * 4 + bn) * 2;
                    const int of = (offw[k >> 2] >> ((k & 3) * 8)) & 0xff;
                    const int op = (offw[(k + 1) >> 2] >> (((k + 1) & 3) * 8)) & 0xff;
                    int cf = 0, cp = 0;
                    for (int db = 0; db < DIM; ++db) {
                        const int c = row_dof[bn][db];
                        const int64_t pos = c < a.n_free ? base + of + cf++ : base + op + cp++;
                        kadd<KT>(a, pos, Kr[bn * DIM + db]);
                    }
                }
            }
    }
}

// ---------------------------------------------------------------------------
// Cut element computed by one wavefront: K_e (M x M) entries handed to
// emit(i, j, value), P_e (M) into LDS.
// Every integration point is staged in LDS first (gradients, J*W, the point's
// material, its stress; per cohesive point the rotated tangent R^T D_c R, the
// rotated traction and the surface shape functions), then each lane owns one
// row of K_e (or a column group of it: entries j = grp + t * nsplit), forms
// that row's D B column once per point and sums every point's term into its
// entries in point order (volumetric points, then cohesive points): three
// barriers per element.  In 2D (M <= 24, at most 12 entries per lane, no
// cohesive points on the path's 2D workloads) the entries live in registers,
// so no LDS tile for K_e limits how many elements share a CU; in 3D (up to 36
// entries per lane, cohesive points matched per entry) they live in an LDS
// row of stride M + 1.  The cohesive law runs once per cohesive point (lane
// q), which also writes that point's trial delta_max and damage flags.
template <int DIM>
__host__ __device__ constexpr int cut_row_entries() {
    return (MAX_NT * DIM + cut_threads<DIM>() / (MAX_NT * DIM) - 1) / (cut_threads<DIM>() / (MAX_NT * DIM));
}

template <int DIM, bool DO_K, bool DO_F, typename Emit>
__device__ inline void cut_compute(const AsmArgs& a, int c, int nt, double* s_P, int* s_node, int* s_dof, Emit emit) {
    const int tid = threadIdx.x;
    constexpr int S = DIM == 2 ? 3 : 6;
    constexpr int MM = MAX_NT * DIM;
    constexpr int NV = max_vpt(DIM);
    const int M = nt * DIM;
    __shared__ double s_u[MM];
    __shared__ double s_g[NV * MM];
    __shared__ double s_w[NV];
    __shared__ int s_vmat[NV];
    __shared__ double s_sig[NV * S];
    __shared__ double s_cRDR[MAX_CPT * DIM * DIM], s_ctR[MAX_CPT * DIM], s_cN[MAX_CPT * 4], s_cw[MAX_CPT];
    __shared__ int s_cmap[MAX_CPT * 8], s_cnh[MAX_CPT];
    const int64_t v0 = a.vpt_ptr[c];
    const int nv = (int)(a.vpt_ptr[c + 1] - v0);
    const int64_t c0 = a.cpt_ptr[c];
    const int nc = (int)(a.cpt_ptr[c + 1] - c0);
    if (tid < nt) s_node[tid] = a.ce_nodes[c * 16 + tid];
    __syncthreads();
    for (int i = tid; i < M; i += cut_threads<DIM>()) {
        s_dof[i] = a.dof[(i % DIM) * a.n_all + s_node[i / DIM]];
        s_u[i] = a.U[(i % DIM) * a.n_all + s_node[i / DIM]];
    }
    for (int k = tid; k < nv * M; k += cut_threads<DIM>()) {
        const int q = k / M;
        s_g[q * MM + (k - q * M)] = a.vpt_grad[a.vpt_grad_off[v0 + q] + (k - q * M)];
    }
    if (tid < nv) {
        s_w[tid] = a.vpt_w[v0 + tid];
        s_vmat[tid] = a.vpt_mat[v0 + tid];
    }
    __syncthreads();
    if (DO_F && tid < nv) {
        double eps[S];
        strain_of<DIM>(nt, &s_g[tid * MM], s_u, eps);
        stress_of<DIM>(iso_of<DIM>(a.Dmat + s_vmat[tid] * 36), eps, &s_sig[tid * S]);
    }
    if (tid < nc) {
        const int64_t q = c0 + tid;
        const int nh = a.cpt_nh[q];
        double R[9], dl[3] = {0, 0, 0}, delta[3] = {0, 0, 0};
        for (int k = 0; k < 9; ++k) R[k] = a.cpt_R[q * 9 + k];
        // jump across the interface: surface node n (lower slot map[n], upper
        // slot map[nh + n]) read straight from the tables, no indexed arrays
        for (int n = 0; n < 4; ++n) {
            if (n >= nh) break;
            const double Nn = a.cpt_N[q * 4 + n];
            const int lo = a.cpt_nodes[q * 8 + n], hi = a.cpt_nodes[q * 8 + nh + n];
            for (int d = 0; d < DIM; ++d) dl[d] += Nn * (s_u[hi * DIM + d] - s_u[lo * DIM + d]);
        }
        for (int i = 0; i < DIM; ++i)
            for (int j = 0; j < DIM; ++j) delta[i] += R[i * DIM + j] * dl[j];
        double trac[3], Dc[9], dnew;
        uint8_t fl;
        coh_law<DIM>(a.laws[a.cpt_law[q]], delta, a.dmax_conv[q], a.dt, trac, Dc, dnew, fl);
        a.dmax_trial[q] = dnew;
        a.flags[q] = fl;
        double RtD[9];
        for (int i = 0; i < DIM; ++i)
            for (int j = 0; j < DIM; ++j) {
                double s = 0.0;
                for (int k = 0; k < DIM; ++k) s += R[k * DIM + i] * Dc[k * DIM + j];
                RtD[i * DIM + j] = s;
            }
        for (int i = 0; i < DIM; ++i)
            for (int j = 0; j < DIM; ++j) {
                double s = 0.0;
                for (int k = 0; k < DIM; ++k) s += RtD[i * DIM + k] * R[k * DIM + j];
                s_cRDR[tid * DIM * DIM + i * DIM + j] = s;
            }
        for (int i = 0; i < DIM; ++i) {
            double s = 0.0;
            for (int k = 0; k < DIM; ++k) s += R[k * DIM + i] * trac[k];
            s_ctR[tid * DIM + i] = s;
        }
        for (int k = 0; k < 4; ++k) s_cN[tid * 4 + k] = a.cpt_N[q * 4 + k];
        for (int k = 0; k < 8; ++k) s_cmap[tid * 8 + k] = k < 2 * nh ? (int)a.cpt_nodes[q * 8 + k] : -1;
        s_cnh[tid] = nh;
        s_cw[tid] = a.cpt_w[q];
    }
    __syncthreads();
    if (DO_K && DIM == 2) {
        // lane -> (row i, column group grp): the lane owns K_e(i, j) for
        // j = grp + t * nsplit, held in registers
        constexpr int TMAX = cut_row_entries<DIM>();
        const int nsplit = cut_threads<DIM>() / M;
        if (tid < M * nsplit) {
            const int i = tid % M, grp = tid / M;
            const int an = i / DIM, da = i % DIM;
            double acc[TMAX];
#pragma unroll
            for (int t = 0; t < TMAX; ++t) acc[t] = 0.0;
            for (int q = 0; q < nv; ++q) {
                double parts[S];
                k_parts_col<DIM>(&s_g[q * MM + an * DIM], da, iso_of<DIM>(a.Dmat + s_vmat[q] * 36), parts);
                const double wq = s_w[q];
                const double* gq = &s_g[q * MM];
#pragma unroll
                for (int t = 0; t < TMAX; ++t) {
                    const int j = grp + t * nsplit;
                    if (j < M) acc[t] += wq * k_entry<DIM>(parts, gq + (j / DIM) * DIM, j % DIM);
                }
            }
            for (int q = 0; q < nc; ++q) {
                const int nh = s_cnh[q];
                for (int k1 = 0; k1 < 2 * nh; ++k1) {
                    if (s_cmap[q * 8 + k1] != an) continue;
                    const double n1 = s_cN[q * 4 + k1 % nh];
#pragma unroll
                    for (int t = 0; t < TMAX; ++t) {
                        const int j = grp + t * nsplit;
                        if (j >= M) continue;
                        const int bn = j / DIM, db = j % DIM;
                        for (int k2 = 0; k2 < 2 * nh; ++k2) {
                            if (s_cmap[q * 8 + k2] != bn) continue;
                            const double sgn = ((k1 < nh) == (k2 < nh)) ? 1.0 : -1.0;
                            const double v = n1 * s_cN[q * 4 + k2 % nh] * s_cRDR[q * DIM * DIM + da * DIM + db];
                            acc[t] += s_cw[q] * (sgn * v);
                        }
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < TMAX; ++t) {
                const int j = grp + t * nsplit;
                if (j < M) emit(i, j, acc[t]);
            }
        }
    }
    if (DO_K && DIM == 3) {
        // lane -> (row i, column group grp): the lane's entries of row i live
        // in LDS (row stride M + 1: lanes of different rows hit different banks)
        __shared__ double s_K[MM * (MM + 1)];
        const int nsplit = cut_threads<DIM>() / M;
        if (tid < M * nsplit) {
            const int i = tid % M, grp = tid / M;
            const int an = i / DIM, da = i % DIM;
            double* Ki = s_K + i * (M + 1);
            for (int j = grp; j < M; j += nsplit) Ki[j] = 0.0;
            for (int q = 0; q < nv; ++q) {
                double parts[S];
                k_parts_col<DIM>(&s_g[q * MM + an * DIM], da, iso_of<DIM>(a.Dmat + s_vmat[q] * 36), parts);
                const double wq = s_w[q];
                const double* gq = &s_g[q * MM];
                for (int j = grp; j < M; j += nsplit) Ki[j] += wq * k_entry<DIM>(parts, gq + (j / DIM) * DIM, j % DIM);
            }
            for (int q = 0; q < nc; ++q) {
                const int nh = s_cnh[q];
                for (int k1 = 0; k1 < 2 * nh; ++k1) {
                    if (s_cmap[q * 8 + k1] != an) continue;
                    for (int k2 = 0; k2 < 2 * nh; ++k2) {
                        const int bn = s_cmap[q * 8 + k2];
                        const double sgn = ((k1 < nh) == (k2 < nh)) ? 1.0 : -1.0;
                        for (int db = 0; db < DIM; ++db) {
                            const int j = bn * DIM + db;
                            if (j % nsplit != grp) continue;
                            const double v = s_cN[q * 4 + k1 % nh] * s_cN[q * 4 + k2 % nh] * s_cRDR[q * DIM * DIM + da * DIM + db];
                            Ki[j] += s_cw[q] * (sgn * v);
                        }
                    }
                }
            }
            for (int j = grp; j < M; j += nsplit) emit(i, j, Ki[j]);
        }
    }
    if (DO_F)
        for (int i = tid; i < M; i += cut_threads<DIM>()) {
            const int an = i / DIM, d = i % DIM;
            double acc = 0.0;
            #pragma unroll 1
            for (int q = 0; q < nv; ++q) acc += s_w[q] * force_entry<DIM>(&s_g[q * MM + an * DIM], d, &s_sig[q * S]);
            #pragma unroll 1
            for (int q = 0; q < nc; ++q) {
                const int nh = s_cnh[q];
                #pragma unroll 1
                for (int k = 0; k < 2 * nh; ++k)
                    if (s_cmap[q * 8 + k] == an) {
                        const double v = s_cN[q * 4 + k % nh] * s_ctR[q * DIM + d];
                        acc += s_cw[q] * (k < nh ? -v : v);
                    }
            }
            s_P[i] = acc;
        }
    __syncthreads();
}

// Coloured scheme, cut elements: compute then scatter into the CSR rows.
template <int DIM, bool DO_K, bool DO_F, typename KT>
__global__ __launch_bounds__(cut_threads<DIM>()) void k_cut(AsmArgs a, int64_t begin, int64_t end) {
    const int64_t li = begin + blockIdx.x;
    if (li >= end) return;
    constexpr int MM = MAX_NT * DIM;
    const int c = a.ce_list[li];
    const int nt = a.ce_nt[c], M = nt * DIM, tid = threadIdx.x;
    __shared__ double s_P[MM];
    __shared__ int s_node[MAX_NT], s_dof[MM];
    const uint8_t* off = a.ce_off + a.ce_off_ptr[c];
    // entries scattered straight from the lanes' registers (one colour: no
    // other element touches these rows' entries meanwhile)
    cut_compute<DIM, DO_K, DO_F>(a, c, nt, s_P, s_node, s_dof, [&](int i, int j, double v) {
        const int r = s_dof[i];
        if (r >= a.n_free && !a.all_rows) return;
        const int an = i / DIM, bn = j / DIM, db = j % DIM;
        const int of = off[(an * nt + bn) * 2], op = off[(an * nt + bn) * 2 + 1];
        int cf = 0, cp = 0;
        for (int k = 0; k < db; ++k) { if (s_dof[bn * DIM + k] < a.n_free) ++cf; else ++cp; }
        const int cdof = s_dof[j];
        const int64_t pos = cdof < a.n_free ? a.rowptr[r] + of + cf : a.rowptr[r] + op + cp;
        kadd<KT>(a, pos, v);
    });
    if (DO_F)
        for (int i = tid; i < M; i += cut_threads<DIM>()) {
            const int r = s_dof[i];
            if (r < a.n_free || a.all_rows) a.P[r] += s_P[i];
        }
}

// ---------------------------------------------------------------------------
// Cut-element matrices on the FP64 matrix cores.
//
// The element matrix of a cut element (src/Assembler.cpp:392-524: volumetric
// sub-element points, then cohesive points) is one small dense product
//     K_e = A^T C,
// with the rows of A and C stacked over the element's integration points:
//   * volumetric point q: with the isotropic D = lam m m^T + mu D0 (D0 =
//     diag(2, 2, 1) in 2D, diag(2, 2, 2, 1, 1, 1) in 3D), A = (m^T B_q; B_q)
//     and C = diag(w lam, 2 w mu (normal rows), w mu (shear rows)) A, so
//     A^T C = w B^T D B (SmallStrainKFast, src/Utilities.cpp:196-296); 4 rows
//     in 2D, 7 in 3D;
//   * cohesive point p (DIM rows, jump component d):  A = L_p,  C = w_p T_p L_p
//     with L_p (u) = sum_n N_n (u_upper,n - u_lower,n) and T_p = R^T D_c R
//     (KernelCohesive / cohesiveKFast, src/AssemblerSmallStrainNonlinear.cpp:73-91,
//     src/Utilities.cpp:480-508: the +-N_i N_j R^T D R blocks).
// One wavefront per cut element, four per block.  Each lane forms the A and C
// operand values it feeds to v_mfma_f64_16x16x4f64 straight from the point data
// staged in LDS (gradients, weights, isotropic D, R^T D_c R, signed surface
// shape functions), so no operand matrix is materialised; K_e (up to 16 x 16
// in 2D, one output tile; up to 36 x 36 in 3D, 2 x 2 or 3 x 3 tiles of
// 16 x 16) accumulates in registers and is written once to the cut-element
// scratch the gather pass reads.  P_e and the cohesive law (trial delta_max
// and damage flags) are as in cut_compute.
//
// v_mfma_f64_16x16x4f64 lane layout (wave64): A operand lane l holds
// A(row l % 16, k l / 16); B operand lane l holds B(k l / 16, col l % 16); the
// four results of lane l are D(4 r + l / 16, l % 16), r = 0..3 (rows
// interleaved across the four lane groups, unlike the fp32 16x16x4 form).
typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int MF_WAVES = 4;  // cut elements per block

template <int DIM>
struct MfWave {
    static constexpr int NV = max_vpt(DIM);  // volumetric points staged
    static constexpr int MM = MAX_NT * DIM;
    static constexpr int S = DIM == 2 ? 3 : 6;
    double g[NV * MM];            // slot gradients per point
    double w[NV];
    double Dabc[NV * 3];          // isotropic D: a (normal), b (off-diagonal), c (shear)
    double wl[NV], wm[NV];        // w lam and w mu of every point (lam = b, mu = c)
    double sig[NV * S];
    double u[MM];
    double T[MAX_CPT * DIM * DIM];  // w_p R^T D_c R
    double cs[MAX_CPT * MAX_NT];    // signed surface shape function of every slot (upper +N, lower -N)
    double ctR[MAX_CPT * DIM];      // w_p R^T t
    int node[MAX_NT];
    int nv, nc, nt;
};

// Operand rows of a volumetric point q for the isotropic D = lam m m^T + mu D0:
// B^T D B = lam (m^T B)^T (m^T B) + mu B^T D0 B, so with the rows
//   s = 0: (m^T B)(a, da) = g_a[da]                      weight w lam
//   3D: s = 1..3 normal Voigt rows, s = 4..6 shear rows yz, xz, xy
//   2D: s = 1..2 normal Voigt rows xx, yy, s = 3 the shear row xy
// (normal rows weight 2 w mu, shear rows w mu) C is the weight times A, and
// an entry of A is one gradient component or 0.
// A(kk, (a, da)) for row s: the gradient component it selects, or -1 for 0.
template <int DIM>
__device__ inline int vrow_comp(int s, int da) {
    if (s == 0) return da;
    if (DIM == 2) {
        if (s <= 2) return s - 1 == da ? da : -1;
        return 1 - da;  // xy: u_x <- g_y, u_y <- g_x
    }
    if (s <= 3) return s - 1 == da ? da : -1;
    if (s == 4) return da == 0 ? -1 : 3 - da;  // yz: u_y <- g_z, u_z <- g_y
    if (s == 5) return da == 1 ? -1 : 2 - da;  // xz: u_x <- g_z, u_z <- g_x
    return da == 2 ? -1 : 1 - da;              // xy: u_x <- g_y, u_y <- g_x
}
template <int DIM>
constexpr int vrows() { return DIM == 2 ? 4 : 7; }

// K_e = A^T C over the stacked rows: vrows per volumetric point, then DIM per
// cohesive point (A = d-th jump row, C = T_p rows times the signed shape
// functions).  Each lane evaluates, per k-step, the A entries of its column
// in every tile (NT values) and scales them into C for volumetric rows.
template <int DIM, int NT>
__device__ inline void mfma_ke(const MfWave<DIM>& W, int M, int lane, double* ke) {
    constexpr int MM = MfWave<DIM>::MM, R = vrows<DIM>();
    v4d acc[NT][NT];
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj) acc[ti][tj] = v4d{0.0, 0.0, 0.0, 0.0};
    const int KV = R * W.nv, KD = KV + DIM * W.nc;
    const int r16 = lane & 15, k4 = lane >> 4;
    // this lane's column of every tile: slot sl, component da
    int sl[NT], da[NT];
    bool ok[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int i = 16 * t + r16;
        ok[t] = i < M;
        sl[t] = i / DIM;
        da[t] = i - DIM * sl[t];
    }
    for (int k0 = 0; k0 < KD; k0 += 4) {
        const int kk = k0 + k4;
        double av[NT], cv[NT];
        if (kk < KV) {
            const int q = kk / R, s = kk - R * q;
            const double wt = s == 0 ? W.wl[q] : (s <= DIM ? 2.0 * W.wm[q] : W.wm[q]);
            const double* gq = &W.g[q * MM];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int cmp = vrow_comp<DIM>(s, da[t]);
                av[t] = ok[t] && cmp >= 0 ? gq[sl[t] * DIM + cmp] : 0.0;
                cv[t] = wt * av[t];
            }
        } else if (kk < KD) {
            const int kc = kk - KV, p = kc / DIM, d = kc - DIM * p;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const double sa = ok[t] ? W.cs[p * MAX_NT + sl[t]] : 0.0;
                av[t] = d == da[t] ? sa : 0.0;
                cv[t] = W.T[p * DIM * DIM + d * DIM + da[t]] * sa;
            }
        } else {
#pragma unroll
            for (int t = 0; t < NT; ++t) av[t] = cv[t] = 0.0;
        }
        // One output tile: each product accumulates into the tile the previous
        // product (about twenty instructions earlier) is still writing, and the
        // compiler inserts no wait for an exactly overlapping accumulator.  On
        // gfx950 that gave wrong K_e that changed from run to run; 16 wait
        // states before each product (placed after the operands are formed)
        // give stable results within the oracle's bar.  With 2 x 2 or 3 x 3
        // tiles the other tiles' products separate two uses of one tile.
        if constexpr (NT == 1) asm volatile("s_nop 7\n\ts_nop 7" : "+v"(av[0]), "+v"(cv[0]));
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = 0; tj < NT; ++tj)
                acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ti], cv[tj], acc[ti][tj], 0, 0, 0);
    }
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * ti + 4 * r + k4, col = 16 * tj + r16;
                if (row < M && col < M)
                    ke[(row / DIM) * cut_slot_doubles(M, DIM) + (row % DIM) * cut_ke_stride(M) + col] = acc[ti][tj][r];
            }
}

// At least four waves per SIMD: the compiler then keeps the 3 x 3 tiles of
// accumulators in ordinary VGPRs (128 in all, no spills) instead of 131 VGPRs
// plus 72 accumulation registers at two waves per SIMD; the staging loads of
// one element overlap the products of the others.
template <int DIM, bool DO_K, bool DO_F>
__global__ __launch_bounds__(MF_WAVES * WAVE, 4) void k_cut_elem_mfma(AsmArgs a) {
    constexpr int MM = MfWave<DIM>::MM, S = MfWave<DIM>::S, SS = S * S;
    __shared__ MfWave<DIM> s_w[MF_WAVES];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int64_t blk = xcd_logical(blockIdx.x, gridDim.x);
    const int64_t cl = blk * MF_WAVES + wave;
    const bool live = cl < a.n_cut;  // padding waves keep to the barriers
    MfWave<DIM>& W = s_w[wave];
    const int c = live ? (int)cl : 0;
    const int nt = live ? a.ce_nt[c] : 0, M = DIM * nt;
    const int64_t v0 = live ? a.vpt_ptr[c] : 0, c0 = live ? a.cpt_ptr[c] : 0;
    const int nv = live ? (int)(a.vpt_ptr[c + 1] - v0) : 0;
    const int nc = live ? (int)(a.cpt_ptr[c + 1] - c0) : 0;
    if (lane == 0) { W.nv = nv; W.nc = nc; W.nt = nt; }
    if (lane < nt) W.node[lane] = a.ce_nodes[c * 16 + lane];
    __syncthreads();
    for (int i = lane; i < M; i += WAVE) W.u[i] = a.U[(i % DIM) * a.n_all + W.node[i / DIM]];
    if (a.grad_contig) {
        // one run of nv * M gradients: every load issued before the first is
        // used (one memory latency, not one per pass of the wave)
        constexpr int GL = (MfWave<DIM>::NV * MM + WAVE - 1) / WAVE;
        const double* gs = a.vpt_grad + (live ? a.vpt_grad_off[v0] : 0);
        double gv[GL];
#pragma unroll
        for (int i = 0; i < GL; ++i) {
            const int k = lane + i * WAVE;
            if (k < nv * M) gv[i] = gs[k];
        }
#pragma unroll
        for (int i = 0; i < GL; ++i) {
            const int k = lane + i * WAVE;
            if (k < nv * M) {
                const int q = k / M;
                W.g[q * MM + (k - q * M)] = gv[i];
            }
        }
    } else {
        for (int k = lane; k < nv * M; k += WAVE) {
            const int q = k / M;
            W.g[q * MM + (k - q * M)] = a.vpt_grad[a.vpt_grad_off[v0 + q] + (k - q * M)];
        }
    }
    if (lane < nv) {
        const double w = a.vpt_w[v0 + lane];
        W.w[lane] = w;
        const double* D = a.Dmat + a.vpt_mat[v0 + lane] * 36;
        W.Dabc[3 * lane] = D[0];
        W.Dabc[3 * lane + 1] = D[1];
        W.Dabc[3 * lane + 2] = D[SS - 1];
        W.wl[lane] = w * D[1];
        W.wm[lane] = w * D[SS - 1];
    }
    for (int k = lane; k < nc * MAX_NT; k += WAVE) W.cs[k] = 0.0;
    __syncthreads();
    if (DO_F && lane < nv) {
        double eps[S];
        strain_of<DIM>(nt, &W.g[lane * MM], W.u, eps);
        const Iso m{W.Dabc[3 * lane], W.Dabc[3 * lane + 1], W.Dabc[3 * lane + 2]};
        stress_of<DIM>(m, eps, &W.sig[lane * S]);
    }
    if (lane < nc) {
        const int64_t q = c0 + lane;
        const int nh = a.cpt_nh[q];
        double R[9], dl[3] = {0, 0, 0}, delta[3] = {0, 0, 0};
        for (int k = 0; k < 9; ++k) R[k] = a.cpt_R[q * 9 + k];
        for (int n = 0; n < 4; ++n) {
            if (n >= nh) break;
            const double Nn = a.cpt_N[q * 4 + n];
            const int lo = a.cpt_nodes[q * 8 + n], hi = a.cpt_nodes[q * 8 + nh + n];
            for (int d = 0; d < DIM; ++d) dl[d] += Nn * (W.u[hi * DIM + d] - W.u[lo * DIM + d]);
            W.cs[lane * MAX_NT + lo] = -Nn;
            W.cs[lane * MAX_NT + hi] = Nn;
        }
        for (int i = 0; i < DIM; ++i)
            for (int j = 0; j < DIM; ++j) delta[i] += R[i * DIM + j] * dl[j];
        double trac[3], Dc[9], dnew;
        uint8_t fl;
        coh_law<DIM>(a.laws[a.cpt_law[q]], delta, a.dmax_conv[q], a.dt, trac, Dc, dnew, fl);
        a.dmax_trial[q] = dnew;
        a.flags[q] = fl;
        const double wq = a.cpt_w[q];
        double RtD[9];
        for (int i = 0; i < DIM; ++i)
            for (int j = 0; j < DIM; ++j) {
                double s = 0.0;
                for (int k = 0; k < DIM; ++k) s += R[k * DIM + i] * Dc[k * DIM + j];
                RtD[i * DIM + j] = s;
            }
        for (int i = 0; i < DIM; ++i)
            for (int j = 0; j < DIM; ++j) {
                double s = 0.0;
                for (int k = 0; k < DIM; ++k) s += RtD[i * DIM + k] * R[k * DIM + j];
                W.T[lane * DIM * DIM + i * DIM + j] = wq * s;
            }
        for (int i = 0; i < DIM; ++i) {
            double s = 0.0;
            for (int k = 0; k < DIM; ++k) s += R[k * DIM + i] * trac[k];
            W.ctR[lane * DIM + i] = wq * s;
        }
    }
    __syncthreads();
    if (live && DO_K) {
        double* ke = a.ke + a.ke_ptr[c];
        if (M <= 16) mfma_ke<DIM, 1>(W, M, lane, ke);
        else if (M <= 32) mfma_ke<DIM, 2>(W, M, lane, ke);
        else mfma_ke<DIM, 3>(W, M, lane, ke);
    }
    if (live && DO_F) {
        // P_e entries of slot an follow its K_e rows and column bytes
        double* pe = a.ke + a.ke_ptr[c];
        for (int i = lane; i < M; i += WAVE) {
            const int an = i / DIM, d = i - DIM * an;
            double acc = 0.0;
#pragma unroll 1
            for (int q = 0; q < nv; ++q) acc += W.w[q] * force_entry<DIM>(&W.g[q * MM + an * DIM], d, &W.sig[q * S]);
#pragma unroll 1
            for (int p = 0; p < nc; ++p) acc += W.cs[p * MAX_NT + an] * W.ctR[p * DIM + d];
            pe[an * cut_slot_doubles(M, DIM) + DIM * cut_ke_stride(M) + cut_rel_doubles(M) + d] = acc;
        }
    }
}


// Adds v to an entry of the block's LDS segment without waiting for the
// result (ds_add_f64, no return).  One lane owns a row in a round, so the adds
// to an entry still happen in the owner's program order (element order) and
// round like a read-add-write, but successive adds of a row (12 per uncut
// element, up to 36 per cut element) no longer wait on each other's LDS round
// trip, which a read-add-write must because two of a row's columns could be
// the same entry for all the compiler knows.
__device__ inline void lds_add(double* p, double v) {
    (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One integration point's contribution to a unit's rows: the K entries
// w B_a^T D B_b through the reference's parts (SmallStrainKFast,
// src/Utilities.cpp:196-296) and the internal force w B_a^T sigma
// (src/Utilities.cpp:366-408).  g holds the gradients of the element's four
// nodes, la is the row node's slot.
template <int DIM, int NR, bool DO_K, bool DO_F>
__device__ inline void qp_rows(const double* g, double w, const double* sig, int la, const int* da, const Iso& D,
                               double (*Kr)[4 * DIM], double* Pc) {
    constexpr int S = DIM == 2 ? 3 : 6;
    // the row node's gradient, selected without indexing the register array
    // by the run-time slot
    double gr[DIM];
    for (int i = 0; i < DIM; ++i)
        gr[i] = la == 0 ? g[i] : la == 1 ? g[DIM + i] : la == 2 ? g[2 * DIM + i] : g[3 * DIM + i];
    for (int h = 0; h < NR; ++h) {
        if (DO_F) Pc[h] += w * force_entry<DIM>(gr, da[h], sig);
        if (DO_K) {
            double parts[S];
            k_parts<DIM>(gr, da[h], D, parts);
            for (int bn = 0; bn < 4; ++bn)
                for (int db = 0; db < DIM; ++db)
                    Kr[h][bn * DIM + db] += w * k_entry<DIM>(parts, &g[bn * DIM], db);
        }
    }
}

// Gather scheme, pass 2.  A block owns a run of consecutive nodes.  Its
// threads are (unit, slot) pairs: a unit is one matrix row (NR = 1) or all DIM
// rows of one node (NR = DIM, the element geometry and the nodal data then
// serve the node's rows together), and the SL slots of a unit (1, 2, 4 or 8)
// take the unit's incident elements s, s + SL, ... so the element work of one
// node runs in parallel.  Rows, slots and rows per thread are a per-context
// choice; every choice gives the same bits.  When the free rows of the block
// form one contiguous CSR segment [s0, s1) it is accumulated in LDS and
// written back with coalesced stores; otherwise (s0 == s1), and for
// prescribed rows (all_rows only), the rows are accumulated in place in HBM.
// The slots add their element's contribution in slot order, so every entry
// sums its incident elements in the node's incidence order (its uncut
// elements in element order, then its cut elements in element order) and the
// result does not depend on the layout or on scheduling.
template <int DIM, int SL, int NR, bool DO_K, bool DO_F, typename KT>
__global__ __launch_bounds__(GATHER_THREADS) void k_gather(AsmArgs a) {
    HIP_DYNAMIC_SHARED(double, s_seg)
    constexpr int NU = GATHER_THREADS / SL;  // units per block
    __shared__ double s_P[NU * NR];
    const int64_t blk = a.gblk_lo + xcd_logical(blockIdx.x, gridDim.x);
    if (blk >= a.n_gblk) return;  // padding block of the XCD mapping
    const int64_t n0 = a.blk_node[blk], n1 = a.blk_node[blk + 1];
    const int64_t s0 = a.blk_seg[2 * blk], s1 = a.blk_seg[2 * blk + 1];
    const int rounds = a.blk_rounds[blk];
    const bool use_lds = DO_K && s1 > s0;
    // The slots of a unit are consecutive lanes of one wavefront and a unit's
    // rows are touched by its own lanes only.  When every row the block sums
    // lives in LDS, the slot order needs no block barrier: the wavefront
    // issues slot 0's LDS operations before slot 1's and LDS executes one
    // wavefront's instructions in order, so each wavefront runs its rounds
    // without waiting for the block's slowest.  Rows summed in HBM (prescribed
    // rows, blocks without a contiguous segment) keep the block barrier.
    const bool block_sync = DO_K && (!use_lds || a.all_rows);
    const int t = threadIdx.x;
    const int lu = t / SL, slot = t % SL;  // local unit, slot
    const int64_t n = NR == 1 ? n0 + lu / DIM : n0 + lu;
    // a periodic image owns no rows (its incidence list is empty; the node
    // holding its dofs first gathers its elements)
    const bool active = n < n1 && !(a.image && a.image[n]);
    constexpr int S = DIM == 2 ? 3 : 6;
    constexpr int NQ = DIM == 2 ? 4 : 1;
    int da[NR], r[NR];
    bool own[NR], in_lds[NR];
    int64_t base[NR];
    bool any = false;
    for (int h = 0; h < NR; ++h) {
        da[h] = NR == 1 ? lu % DIM : h;
        r[h] = active ? a.dof[da[h] * a.n_all + n] : 0;
        own[h] = active && (r[h] < a.n_free || a.all_rows);
        in_lds[h] = use_lds && r[h] < a.n_free;
        base[h] = own[h] ? a.rowptr[r[h]] : 0;
        any = any || own[h];
    }
    if (use_lds)
        for (int64_t j = t; j < s1 - s0; j += GATHER_THREADS) s_seg[j] = 0.0;
    // rows outside the LDS segment are summed in FP64 in HBM (the FP64 array
    // is that sum's scratch in FP32 mode), so an FP32 entry is always the
    // FP64 sum rounded once, whatever the layout
    if (DO_K && slot == 0)
        for (int h = 0; h < NR; ++h)
            if (own[h] && !in_lds[h])
                for (int32_t j = a.rowptr[r[h]]; j < a.rowptr[r[h] + 1]; ++j) a.K[j] = 0.0;
    if (DO_F && slot == 0)
        for (int h = 0; h < NR; ++h) s_P[lu * NR + h] = 0.0;
    __syncthreads();
    const int64_t kbeg = any ? a.inc_ptr[n] : 0, kend = any ? a.inc_ptr[n + 1] : 0;
    for (int rr = 0; rr < rounds; ++rr) {
        const int64_t k = kbeg + (int64_t)rr * SL + slot;
        const bool valid = any && k < kend;
        // ---- compute: this slot's element contribution to the unit's rows ----
        int nd[4];
        double Pc[NR];
        for (int h = 0; h < NR; ++h) Pc[h] = 0.0;
        int c = -1, la = 0;
        // One unit row (NR == 1): the row's entries from this incidence and
        // their column bytes, in one buffer for both element kinds (an uncut
        // element's 4 * DIM entries, or a cut element's K_e row loaded in one
        // batch: one memory latency per cut incidence instead of one per
        // entry), so both kinds share one add loop and one set of registers.
        // The buffer holds the first CH entries (20 in 3D, 8 in 2D: an uncut
        // element's row fits); the rest of a longer cut row is loaded in the
        // add phase, CH entries at a time, which keeps the gather at four
        // waves per SIMD.
        constexpr int MM = MAX_NT * DIM;
        constexpr int CH = DIM == 3 ? 20 : 8;
        static_assert(CH % 4 == 0 && CH >= 4 * DIM && CH <= MM, "row buffer: whole words, an uncut row fits");
        double Kc[NR == 1 ? CH : 1];
        uint32_t rc[NR == 1 ? CH / 4 : 1];
        double Kn[NR == 1 ? 1 : NR][4 * DIM];
        double (*Kr)[4 * DIM] = NR == 1 ? reinterpret_cast<double (*)[4 * DIM]>(Kc) : Kn;
        int Mc = 0;
        int2 rec = {0, 0};
        if (valid) {
            rec = a.inc_rec[k];
            la = rec.y & 0xff;
            c = (rec.y & 0x100) ? 0 : -1;
            if (c < 0) {
                const int mat = rec.y >> 16;
                const Iso D = iso_of<DIM>(a.Dmat + mat * 36);
                for (int h = 0; h < NR; ++h)
                    for (int j = 0; j < 4 * DIM; ++j) Kr[h][j] = 0.0;
                if (DO_K && NR == 1) {
                    // row-relative column positions, one byte per column
                    const uint32_t* rw = (const uint32_t*)(a.inc_rel + k * (4 * DIM));
                    for (int i = 0; i < DIM; ++i) rc[i] = rw[i];
                    Mc = 4 * DIM;
                }
                double X[4][DIM];
                const int4 cn = a.el_conn4[rec.x];
                nd[0] = cn.x; nd[1] = cn.y; nd[2] = cn.z; nd[3] = cn.w;
                for (int q = 0; q < 4; ++q) {
                    if (DIM == 2) {
                        const double2 xv = ((const double2*)a.Xq)[nd[q]];
                        X[q][0] = xv.x;
                        X[q][1] = xv.y;
                    } else {
                        const double4 xv = ((const double4*)a.Xq)[nd[q]];
                        X[q][0] = xv.x;
                        X[q][1] = xv.y;
                        X[q][DIM - 1] = xv.z;
                    }
                }
                double u[4 * DIM];
                if (DO_F)
                    for (int q = 0; q < 4; ++q)
                        for (int d = 0; d < DIM; ++d) u[q * DIM + d] = a.U[d * a.n_all + nd[q]];
                const bool affine = DIM == 2 && quad_affine((const double(*)[2])X);
                if (affine) {
                    quad_affine_rows<NR>((const double(*)[2])X, la, da, D, (double(*)[8])Kr);
                    if (DO_F)
                        for (int h = 0; h < NR; ++h)
                            for (int j = 0; j < 4 * DIM; ++j) Pc[h] += Kr[h][j] * u[j];
                }
                // one quadrature point at a time keeps the register footprint small
#pragma unroll 1
                for (int q = 0; q < (affine ? 0 : NQ); ++q) {
                    double g[4 * DIM], w;
                    qp_geometry<DIM>(X, a.tet_scale, q, g, w);
                    double sig[S];
                    if (DO_F) {
                        double eps[S];
                        strain_of<DIM>(4, g, u, eps);
                        stress_of<DIM>(D, eps, sig);
                    }
                    qp_rows<DIM, NR, DO_K, DO_F>(g, w, sig, la, da, D, Kr, Pc);
                }
            } else {
                // the record holds the scratch offsets of this slot's rows
                Mc = (rec.y >> 9) & 0x7f;
                if (DO_F)
                    for (int h = 0; h < NR; ++h)
                        Pc[h] = a.ke[rec.x + DIM * cut_ke_stride(Mc) + cut_rel_doubles(Mc) + da[h]];
                if (DO_K && NR == 1) {
                    const double2* kr = (const double2*)(a.ke + rec.x + (int64_t)da[0] * cut_ke_stride(Mc));
                    const uint32_t* rw = (const uint32_t*)(a.ke + rec.x + DIM * cut_ke_stride(Mc));
#pragma unroll
                    for (int j = 0; j < CH / 2; ++j)
                        if (2 * j < Mc) {
                            const double2 v = kr[j];
                            Kc[2 * j] = v.x;
                            Kc[2 * j + 1] = v.y;
                        }
#pragma unroll
                    for (int j = 0; j < CH / 4; ++j)
                        if (4 * j < Mc) rc[j] = rw[j];
                }
            }
        }
        // ---- add, slot by slot (element order) ----
        for (int ss = 0; ss < SL; ++ss) {
            if (valid && slot == ss) {
                if (DO_F)
                    for (int h = 0; h < NR; ++h) s_P[lu * NR + h] += Pc[h];
                if (DO_K) {
                    auto add = [&](int h, int64_t rel, double v) {
                        if (in_lds[h]) lds_add(&s_seg[base[h] + rel - s0], v);
                        else a.K[base[h] + rel] += v;
                    };
                    if (NR == 1) {
#pragma unroll
                        for (int j = 0; j < CH; ++j)
                            if (j < Mc && own[0]) add(0, (int)((rc[j / 4] >> (8 * (j % 4))) & 0xffu), Kc[j]);
                        // the rest of a cut element's row, CH entries at a time
                        // (every index stays inside the CH-entry buffer)
                        constexpr int NCH = (MM - 1) / CH;  // further chunks at most
#pragma unroll
                        for (int ch = 1; ch <= NCH; ++ch) {
                            const int j0 = ch * CH;
                            if (j0 >= Mc) break;
                            const double2* kr =
                                (const double2*)(a.ke + rec.x + (int64_t)da[0] * cut_ke_stride(Mc) + j0);
                            const uint32_t* rw =
                                (const uint32_t*)((const uint8_t*)(a.ke + rec.x + DIM * cut_ke_stride(Mc)) + j0);
#pragma unroll
                            for (int j = 0; j < CH / 2; ++j)
                                if (j0 + 2 * j < Mc) {
                                    const double2 v = kr[j];
                                    Kc[2 * j] = v.x;
                                    Kc[2 * j + 1] = v.y;
                                }
#pragma unroll
                            for (int j = 0; j < CH / 4; ++j)
                                if (j0 + 4 * j < Mc) rc[j] = rw[j];
#pragma unroll
                            for (int j = 0; j < CH; ++j)
                                if (j0 + j < Mc && own[0]) add(0, (int)((rc[j / 4] >> (8 * (j % 4))) & 0xffu), Kc[j]);
                        }
                    } else if (c < 0) {
                        // row-relative column positions, one byte per column
                        uint32_t wv[DIM];
                        const uint32_t* rw = (const uint32_t*)(a.inc_rel + k * (4 * DIM));
                        for (int i = 0; i < DIM; ++i) wv[i] = rw[i];
                        for (int j = 0; j < 4 * DIM; ++j) {
                            const int rel = (int)((wv[j / 4] >> (8 * (j % 4))) & 0xffu);
                            for (int h = 0; h < NR; ++h)
                                if (own[h]) add(h, rel, Kr[h][j]);
                        }
                    } else {
                        const uint8_t* rl = (const uint8_t*)(a.ke + rec.x + DIM * cut_ke_stride(Mc));
                        const double* kb = a.ke + rec.x;
                        for (int j = 0; j < Mc; ++j) {
                            const int rel = rl[j];
                            for (int h = 0; h < NR; ++h)
                                if (own[h]) add(h, rel, kb[(int64_t)da[h] * cut_ke_stride(Mc) + j]);
                        }
                    }
                }
            }
            if (SL > 1) {
                if (block_sync) __syncthreads();
                else __builtin_amdgcn_wave_barrier();  // keeps the slots' code in slot order
            }
        }
    }
    if (DO_F && slot == 0)
        for (int h = 0; h < NR; ++h)
            if (own[h]) a.P[r[h]] = s_P[lu * NR + h];
    // (rows summed in HBM: with several slots the last round ended on a block
    // barrier, every add is done)
    if (DO_K && sizeof(KT) == 4 && slot == 0)
        for (int h = 0; h < NR; ++h)
            if (own[h] && !in_lds[h])
                for (int32_t j = a.rowptr[r[h]]; j < a.rowptr[r[h] + 1]; ++j) a.K32[j] = (float)a.K[j];
    if (use_lds) {
        __syncthreads();
        for (int64_t j = t; j < s1 - s0; j += GATHER_THREADS) {
            if (sizeof(KT) == 8) a.K[s0 + j] = s_seg[j];
            else a.K32[s0 + j] = (float)s_seg[j];
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 sums rounded to the FP32 matrix (coloured scheme in FP32 mode)
__global__ void k_round32(int64_t n, const double* __restrict__ src, float* __restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

// ---------------------------------------------------------------------------
// Homogenised response of an RVE (SubscaleModel::homogenize_stress_strain,
// src/SubscaleModel.cpp:42-196): over every volumetric point of the elements
// of a segment, sum w * sigma with sigma = D (eps(u) + eps_bar), w * eps(u)
// and w.  Segment k (elements [seg[k], seg[k+1])) is split evenly over
// VOL_BLOCKS blocks; each block writes one partial per quantity, and
// k_vol_finish adds a segment's partials in block order (deterministic).
constexpr int VOL_BLOCKS = 64;

template <int DIM>
__global__ __launch_bounds__(256) void k_vol_partials(AsmArgs a, const double* __restrict__ eps_bar,
                                                      const int64_t* __restrict__ seg, double* __restrict__ part) {
    constexpr int S = DIM == 2 ? 3 : 6, NQ = DIM == 2 ? 4 : 1, W = 2 * S + 1;
    const int k = blockIdx.x / VOL_BLOCKS, b = blockIdx.x % VOL_BLOCKS;
    const int64_t e0 = seg[k], len = seg[k + 1] - e0;
    const int64_t c0 = e0 + len * b / VOL_BLOCKS, c1 = e0 + len * (b + 1) / VOL_BLOCKS;
    double acc[W];
    for (int i = 0; i < W; ++i) acc[i] = 0.0;
    double eb[S];
    for (int i = 0; i < S; ++i) eb[i] = eps_bar[i];
    auto add_point = [&](const Iso& D, const double* eps, double w) {
        double tot[S], sig[S];
        for (int i = 0; i < S; ++i) tot[i] = eps[i] + eb[i];
        stress_of<DIM>(D, tot, sig);
        for (int i = 0; i < S; ++i) {
            acc[i] += w * sig[i];
            acc[S + i] += w * eps[i];
        }
        acc[2 * S] += w;
    };
    for (int64_t e = c0 + threadIdx.x; e < c1; e += blockDim.x) {
        const int c = a.cut_index[e];
        if (c < 0) {
            double X[4][DIM], u[4 * DIM];
            for (int q = 0; q < 4; ++q) {
                const int nd = a.el_conn[(int64_t)q * a.n_elem + e];
                for (int i = 0; i < DIM; ++i) {
                    X[q][i] = a.X[i * a.n_all + nd];
                    u[q * DIM + i] = a.U[i * a.n_all + nd];
                }
            }
            const Iso D = iso_of<DIM>(a.Dmat + a.el_mat[e] * 36);
            for (int q = 0; q < NQ; ++q) {
                double g[4 * DIM], w, eps[S];
                qp_geometry<DIM>(X, a.tet_scale, q, g, w);
                strain_of<DIM>(4, g, u, eps);
                add_point(D, eps, w);
            }
        } else {
            const int nt = a.ce_nt[c];
            double u[MAX_NT * DIM];
            for (int s = 0; s < nt; ++s) {
                const int nd = a.ce_nodes[c * 16 + s];
                for (int i = 0; i < DIM; ++i) u[s * DIM + i] = a.U[i * a.n_all + nd];
            }
            for (int64_t pt = a.vpt_ptr[c]; pt < a.vpt_ptr[c + 1]; ++pt) {
                double eps[S];
                strain_of<DIM>(nt, a.vpt_grad + a.vpt_grad_off[pt], u, eps);
                add_point(iso_of<DIM>(a.Dmat + a.vpt_mat[pt] * 36), eps, a.vpt_w[pt]);
            }
        }
    }
    // block tree sum in a fixed order
    __shared__ double red[256];
    for (int i = 0; i < W; ++i) {
        red[threadIdx.x] = acc[i];
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * W + i] = red[0];
        __syncthreads();
    }
}

// out[k] = (averaged stress [S], averaged fluctuation strain [S], volume)
__global__ void k_vol_finish(int n_seg, int S, const double* __restrict__ part, double* __restrict__ out) {
    const int W = 2 * S + 1;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n_seg; k += gridDim.x * blockDim.x) {
        double sum[13];
        for (int i = 0; i < W; ++i) sum[i] = 0.0;
        for (int b = 0; b < VOL_BLOCKS; ++b)
            for (int i = 0; i < W; ++i) sum[i] += part[((int64_t)k * VOL_BLOCKS + b) * W + i];
        const double v = sum[2 * S];
        for (int i = 0; i < 2 * S; ++i) out[(int64_t)k * W + i] = v != 0.0 ? sum[i] / v : 0.0;
        out[(int64_t)k * W + 2 * S] = v;
    }
}

// nodal field from a dof vector: every node slot takes its dof's value
// (free part x, prescribed part up), periodic images included
__global__ void k_set_from_dofs(int64_t n_slots, const int32_t* __restrict__ dof, int64_t n_free,
                                const double* __restrict__ x, const double* __restrict__ up, double* __restrict__ U) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_slots; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t g = dof[i];
        U[i] = g < n_free ? x[g] : (up ? up[g - n_free] : 0.0);
    }
}

}  // namespace pgfk

// ---------------------------------------------------------------------------
// Launch one full assembly on the context stream.
namespace pgfd {

template <int DIM, bool DO_K, bool DO_F, typename KT>
static void launch_colors(const Ctx& c, const pgfk::AsmArgs& a) {
    for (int col = 0; col < c.n_colors; ++col) {
        int64_t b = c.ue_color_ptr[col], e = c.ue_color_ptr[col + 1];
        if (e > b) {
            int64_t nb = (e - b + 255) / 256;
            hipLaunchKernelGGL((pgfk::k_uncut<DIM, DO_K, DO_F, KT>), dim3((unsigned)nb), dim3(256), 0, c.stream, a, b, e);
        }
        int64_t cb = c.ce_color_ptr[col], ce = c.ce_color_ptr[col + 1];
        if (ce > cb)
            hipLaunchKernelGGL((pgfk::k_cut<DIM, DO_K, DO_F, KT>), dim3((unsigned)(ce - cb)), dim3(pgfk::cut_threads<DIM>()), 0, c.stream, a,
                               cb, ce);
    }
}

template <int DIM, bool DO_K, bool DO_F, typename KT>
static void launch_gather(const Ctx& c, const pgfk::AsmArgs& a) {
    if (c.timing_events) (void)hipEventRecord(c.timing_events[0], c.stream);
    if (c.n_cut > 0) {
        // four cut elements per block, K_e on the FP64 matrix cores
        const int64_t nb = (c.n_cut + pgfk::MF_WAVES - 1) / pgfk::MF_WAVES;
        hipLaunchKernelGGL((pgfk::k_cut_elem_mfma<DIM, DO_K, DO_F>), dim3((unsigned)pgfk::xcd_blocks(nb)),
                           dim3(pgfk::MF_WAVES * pgfk::WAVE), 0, c.stream, a);
    }
    if (c.timing_events) (void)hipEventRecord(c.timing_events[1], c.stream);
    const size_t lds = DO_K ? sizeof(double) * (size_t)c.seg_max : 0;
    // 3D: regular-node blocks, then enrichment-node blocks (rows filled from
    // the cut elements' K_e only), two launches of the same kernel (128^3 x 6
    // tets with 200 spheres: 6.99 ms against 7.43 ms in one launch); 2D: one
    // launch over all blocks (1000 x 1000 quads: 0.220 ms against 0.236 ms)
    const int64_t split = DIM == 3 ? c.n_gblk_reg : c.n_gblk;
    const int64_t ranges[2][2] = {{0, split}, {split, c.n_gblk}};
    for (const auto& rg : ranges) {
        if (rg[1] <= rg[0]) continue;
        pgfk::AsmArgs ar = a;
        ar.gblk_lo = rg[0];
        ar.n_gblk = rg[1];
        const dim3 grid((unsigned)pgfk::xcd_blocks(rg[1] - rg[0])), blk(pgfk::GATHER_THREADS);
#define PGF_GATHER(SL, NR) hipLaunchKernelGGL((pgfk::k_gather<DIM, SL, NR, DO_K, DO_F, KT>), grid, blk, lds, c.stream, ar)
        if (c.gnode) {
            switch (c.gslots) {
                case 1: PGF_GATHER(1, DIM); break;
                case 2: PGF_GATHER(2, DIM); break;
                case 4: PGF_GATHER(4, DIM); break;
                default: PGF_GATHER(8, DIM); break;
            }
        } else {
            switch (c.gslots) {
                case 1: PGF_GATHER(1, 1); break;
                case 2: PGF_GATHER(2, 1); break;
                case 4: PGF_GATHER(4, 1); break;
                default: PGF_GATHER(8, 1); break;
            }
        }
#undef PGF_GATHER
    }
    if (c.timing_events) (void)hipEventRecord(c.timing_events[2], c.stream);
}

template <int DIM>
static void launch_dim(Ctx& c, const pgfk::AsmArgs& a, bool jac, bool func, bool fp32, bool colored) {
    if (colored) {
        // FP32 mode sums in the FP64 array and rounds each entry once at the end
        if (jac) fill_async(c, c.K, c.nnz, 0.0);
        if (func) fill_async(c, c.P, c.n_rows, 0.0);
        if (jac && func && !fp32) launch_colors<DIM, true, true, double>(c, a);
        else if (jac) launch_colors<DIM, true, false, double>(c, a);
        if (jac && fp32)
            hipLaunchKernelGGL(pgfk::k_round32, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((c.nnz + 255) / 256, 65536))),
                               dim3(256), 0, c.stream, c.nnz, (const double*)c.K, c.K32);
        if (func && (fp32 || !jac)) launch_colors<DIM, false, true, double>(c, a);
        return;
    }
    if (jac && func && !fp32) launch_gather<DIM, true, true, double>(c, a);
    else if (jac && !fp32) launch_gather<DIM, true, false, double>(c, a);
    else if (jac && fp32) launch_gather<DIM, true, false, float>(c, a);
    if (func && (fp32 || !jac)) launch_gather<DIM, false, true, double>(c, a);
}

void enqueue_volume_average(Ctx& c, const double* d_eps_bar, int n_seg, const int64_t* d_seg, double* d_part,
                            double* d_out) {
    const pgfk::AsmArgs a = c.asm_args();
    const dim3 grid((unsigned)(n_seg * pgfk::VOL_BLOCKS)), blk(256);
    if (c.dim == 2) hipLaunchKernelGGL((pgfk::k_vol_partials<2>), grid, blk, 0, c.stream, a, d_eps_bar, d_seg, d_part);
    else hipLaunchKernelGGL((pgfk::k_vol_partials<3>), grid, blk, 0, c.stream, a, d_eps_bar, d_seg, d_part);
    hipLaunchKernelGGL(pgfk::k_vol_finish, dim3((unsigned)((n_seg + 63) / 64)), dim3(64), 0, c.stream, n_seg,
                       c.dim == 2 ? 3 : 6, d_part, d_out);
}

void enqueue_set_from_dofs(Ctx& c, const double* d_x, const double* d_up) {
    const int64_t n = (int64_t)c.dim * c.n_all;
    const unsigned nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(pgfk::k_set_from_dofs, dim3(nb), dim3(256), 0, c.stream, n, c.dof, c.n_free, d_x, d_up, c.U);
}

void enqueue_assembly(Ctx& c, double dt, uint32_t flags) {
    const bool jac = flags & 2u, func = flags & 1u, fp32 = flags & 4u;
    const bool all_rows = flags & 8u, colored = flags & 16u;
    pgfk::AsmArgs a = c.asm_args();
    a.dt = dt;
    a.all_rows = all_rows ? 1 : 0;
    if (c.dim == 2) launch_dim<2>(c, a, jac, func, fp32, colored);
    else launch_dim<3>(c, a, jac, func, fp32, colored);
}

}  // namespace pgfd
