// Shape-sensitivity pseudo-loads on the device.
//
// Per cut element and volumetric integration point, the reference's
// SensitivityShapeParameterStructural::KernelVolumetric
// (src/SensitivityShapeParameterStructural.cpp:18-61) for a linear elastic
// material (no material-parameter stress derivative):
//   dP_e += w * ( B^T (D dB u + sigma div v) + dB^T sigma ),
//   sigma = D (B u + eps_bar),
// where dB is built from the derivative of the point's slot gradients and
// div v from the point's velocity field (host tables, pgf_sens.cpp).  The
// element vectors go to a scratch buffer, then one thread per node adds the
// vectors of its incident cut elements of the selected parameter slot into
// the node's rows in element order (assemble_elem_dPdd + set_vec of
// SensitivityShapeParameter, src/SensitivityShapeParameter.cpp:264-…), so the
// pseudo-load is deterministic and atomic-free.  The integral of the explicit
// stress derivative D dB u + sigma div v is kept per cut element for the
// homogenised-stress sensitivity.
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../../include/pgfem.h"
#include "../host/pgf_common.hpp"
#include "pgf_device.hpp"
#include "pgf_kernels.hpp"

#pragma clang fp contract(on)

namespace pgfk {

template <int DIM>
__device__ inline void sens_strain(int nt, const double* g, const double* u, double* eps) {
    if (DIM == 2) {
        eps[0] = eps[1] = eps[2] = 0.0;
        for (int n = 0; n < nt; ++n) {
            eps[0] += g[n * 2 + 0] * u[n * 2];
            eps[1] += g[n * 2 + 1] * u[n * 2 + 1];
            eps[2] += g[n * 2 + 1] * u[n * 2] + g[n * 2 + 0] * u[n * 2 + 1];
        }
    } else {
        for (int k = 0; k < 6; ++k) eps[k] = 0.0;
        for (int n = 0; n < nt; ++n) {
            eps[0] += g[n * 3 + 0] * u[n * 3];
            eps[1] += g[n * 3 + 1] * u[n * 3 + 1];
            eps[2] += g[n * 3 + 2] * u[n * 3 + 2];
            eps[3] += g[n * 3 + 2] * u[n * 3 + 1] + g[n * 3 + 1] * u[n * 3 + 2];
            eps[4] += g[n * 3 + 2] * u[n * 3] + g[n * 3 + 0] * u[n * 3 + 2];
            eps[5] += g[n * 3 + 1] * u[n * 3] + g[n * 3 + 0] * u[n * 3 + 1];
        }
    }
}
template <int DIM>
__device__ inline void sens_stress(const double* D, const double* e, double* s) {
    // isotropic D: a on the normal diagonal, b off it, c on the shear diagonal
    constexpr int S = DIM == 2 ? 3 : 6;
    const double a = D[0], b = D[1], c = D[S * S - 1];
    if (DIM == 2) {
        s[0] = a * e[0] + b * e[1];
        s[1] = b * e[0] + a * e[1];
        s[2] = c * e[2];
    } else {
        s[0] = a * e[0] + b * e[1] + b * e[2];
        s[1] = b * e[0] + a * e[1] + b * e[2];
        s[2] = b * e[0] + b * e[1] + a * e[2];
        s[3] = c * e[3];
        s[4] = c * e[4];
        s[5] = c * e[5];
    }
}
template <int DIM>
__device__ inline double sens_bts(const double* g, int d, const double* s) {
    if (DIM == 2) return d == 0 ? g[0] * s[0] + g[1] * s[2] : g[1] * s[1] + g[0] * s[2];
    if (d == 0) return g[0] * s[0] + g[2] * s[4] + g[1] * s[5];
    if (d == 1) return g[1] * s[1] + g[2] * s[3] + g[0] * s[5];
    return g[2] * s[2] + g[1] * s[3] + g[0] * s[4];
}

struct SensArgs {
    int64_t n_cut, n_all;
    const int32_t* ce_nodes;   // [n_cut][16]
    const int32_t* ce_nt;
    const int64_t* vpt_ptr;
    const double* vpt_w;
    const int32_t* vpt_mat;
    const int64_t* vpt_grad_off;
    const double* vpt_grad;
    const double* dgrad;       // layout of vpt_grad
    const double* divv;        // [n_vpt]
    const double* Dmat;        // [n_mat][36]
    const double* U;           // [dim][n_all]
    const double* eps_bar;     // [S]
    double* pe;                // [n_cut][MAX_NT * DIM] element pseudo-loads
    double* ds;                // [n_cut][S + 1] explicit stress derivative integrals, then sum of w div v
};

template <int DIM>
__global__ __launch_bounds__(128) void k_sens_elem(SensArgs a) {
    constexpr int S = DIM == 2 ? 3 : 6;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n_cut) return;
    const int nt = a.ce_nt[c], M = nt * DIM;
    double u[MAX_NT * DIM], g[MAX_NT * DIM], eb[S], ds[S], dvol = 0.0;
    for (int i = 0; i < S; ++i) {
        eb[i] = a.eps_bar[i];
        ds[i] = 0.0;
    }
    for (int s = 0; s < nt; ++s) {
        const int nd = a.ce_nodes[c * 16 + s];
        for (int d = 0; d < DIM; ++d) u[s * DIM + d] = a.U[d * a.n_all + nd];
    }
    for (int i = 0; i < M; ++i) g[i] = 0.0;
    for (int64_t pt = a.vpt_ptr[c]; pt < a.vpt_ptr[c + 1]; ++pt) {
        const double* gr = a.vpt_grad + a.vpt_grad_off[pt];
        const double* dg = a.dgrad + a.vpt_grad_off[pt];
        const double* D = a.Dmat + a.vpt_mat[pt] * 36;
        const double w = a.vpt_w[pt], dv = a.divv[pt];
        double eps[S], tot[S], sig[S], deps[S], dsig[S], t[S];
        sens_strain<DIM>(nt, gr, u, eps);
        for (int i = 0; i < S; ++i) tot[i] = eps[i] + eb[i];
        sens_stress<DIM>(D, tot, sig);
        sens_strain<DIM>(nt, dg, u, deps);
        sens_stress<DIM>(D, deps, dsig);
        for (int i = 0; i < S; ++i) {
            t[i] = dsig[i] + sig[i] * dv;
            ds[i] += w * t[i];
        }
        dvol += w * dv;
        for (int s = 0; s < nt; ++s)
            for (int d = 0; d < DIM; ++d)
                g[s * DIM + d] += w * (sens_bts<DIM>(gr + s * DIM, d, t) + sens_bts<DIM>(dg + s * DIM, d, sig));
    }
    for (int i = 0; i < M; ++i) a.pe[c * (MAX_NT * DIM) + i] = g[i];
    for (int i = 0; i < S; ++i) a.ds[c * (S + 1) + i] = ds[i];
    a.ds[c * (S + 1) + S] = dvol;
}

// rhs[dof] = -(sum over the node's incident cut elements of parameter `slot`,
// in element order, of their pseudo-load entries), free dofs; one thread per
// node that owns its rows
template <int DIM>
__global__ __launch_bounds__(256) void k_sens_rows(int64_t n_all, const int32_t* __restrict__ dof, int64_t n_free,
                                                   const uint8_t* __restrict__ image, const int64_t* __restrict__ inc_ptr,
                                                   const int32_t* __restrict__ inc_elem, const uint8_t* __restrict__ inc_slot,
                                                   const int32_t* __restrict__ cut_index, const int32_t* __restrict__ param,
                                                   int32_t slot, const double* __restrict__ pe, double* __restrict__ rhs) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_all || (image && image[n])) return;
    double acc[DIM];
    for (int d = 0; d < DIM; ++d) acc[d] = 0.0;
    for (int64_t k = inc_ptr[n]; k < inc_ptr[n + 1]; ++k) {
        const int c = cut_index[inc_elem[k]];
        if (c < 0 || param[c] != slot) continue;
        const int la = inc_slot[k];
        for (int d = 0; d < DIM; ++d) acc[d] += pe[(int64_t)c * (MAX_NT * DIM) + la * DIM + d];
    }
    for (int d = 0; d < DIM; ++d) {
        const int32_t r = dof[d * n_all + n];
        if (r < n_free) rhs[r] = -acc[d];
    }
}

}  // namespace pgfk

namespace {

int sguard(pgfd::Ctx* c, int (*body)(pgfd::Ctx*, void*), void* user) {
    try {
        return body(c, user);
    } catch (const pgf::Error& e) {
        c->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        return PGF_EDEVICE;
    }
}

#define SCHK(x)                                                                          \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) throw pgf::Error(PGF_EDEVICE, hipGetErrorString(e_));      \
    } while (0)

template <typename T>
T* up(const T* h, int64_t n) {
    T* d = nullptr;
    SCHK(hipMalloc((void**)&d, sizeof(T) * std::max<int64_t>(n, 1)));
    if (n > 0) SCHK(hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice));
    return d;
}

}  // namespace

extern "C" {

PGF_API int pgf_set_shape_sensitivity(pgf_ctx* h, const double* dgrad, const double* divv, const int32_t* param) {
    if (!h || !dgrad || !divv || !param) return PGF_EARG;
    pgfd::Ctx* c = pgfd::ctx_of(h);
    struct A { const double *dg, *dv; const int32_t* pr; } args{dgrad, divv, param};
    return sguard(c, [](pgfd::Ctx* c, void* u) -> int {
        auto* a = (A*)u;
        if (c->n_cpt) throw pgf::Error(PGF_EUNSUPPORTED, "shape sensitivities of cohesive meshes are not built");
        for (void* p : {(void*)c->sens_dgrad, (void*)c->sens_divv, (void*)c->sens_param, (void*)c->sens_pe,
                        (void*)c->sens_ds})
            if (p) SCHK(hipFree(p));
        c->sens_dgrad = up(a->dg, c->n_vgrad);
        c->sens_divv = up(a->dv, c->n_vpt);
        c->sens_param = up(a->pr, c->n_cut);
        const int S = c->dim == 2 ? 3 : 6;
        SCHK(hipMalloc((void**)&c->sens_pe, sizeof(double) * std::max<int64_t>(c->n_cut, 1) * pgfk::MAX_NT * c->dim));
        SCHK(hipMalloc((void**)&c->sens_ds, sizeof(double) * std::max<int64_t>(c->n_cut, 1) * (S + 1)));
        c->h_sens_param.assign(a->pr, a->pr + c->n_cut);
        return 0;
    }, &args);
}

PGF_API int pgf_shape_pseudo_load(pgf_ctx* h, const double* eps_bar, int32_t slot, double* rhs, int32_t n_seg,
                                  const int64_t* seg_elem, double* dsig) {
    if (!h || !eps_bar) return PGF_EARG;
    if (n_seg > 0 && (!seg_elem || !dsig)) return PGF_EARG;
    pgfd::Ctx* c = pgfd::ctx_of(h);
    struct A { const double* eb; int32_t slot; double* rhs; int32_t n_seg; const int64_t* seg; double* dsig; } args{
        eps_bar, slot, rhs, n_seg, seg_elem, dsig};
    return sguard(c, [](pgfd::Ctx* c, void* u) -> int {
        auto* a = (A*)u;
        if (!c->sens_dgrad) throw pgf::Error(PGF_ESTATE, "no shape-sensitivity tables (pgf_set_shape_sensitivity)");
        const int S = c->dim == 2 ? 3 : 6;
        for (int32_t k = 0; k < a->n_seg; ++k)
            if (a->seg[k] < 0 || a->seg[k] > a->seg[k + 1] || a->seg[k + 1] > c->n_elem)
                throw pgf::Error(PGF_EARG, "element segments must be ordered ranges inside [0, n_elem]");
        double* deb = up(a->eb, S);
        double* drhs = nullptr;
        SCHK(hipMalloc((void**)&drhs, sizeof(double) * std::max<int64_t>(c->n_free, 1)));
        pgfk::SensArgs sa{c->n_cut, c->n_all, c->ce_nodes, c->ce_nt, c->vpt_ptr, c->vpt_w, c->vpt_mat,
                          c->vpt_grad_off, c->vpt_grad, c->sens_dgrad, c->sens_divv, c->Dmat, c->U, deb,
                          c->sens_pe, c->sens_ds};
        if (c->n_cut > 0) {
            const unsigned nb = (unsigned)((c->n_cut + 127) / 128);
            if (c->dim == 2) hipLaunchKernelGGL((pgfk::k_sens_elem<2>), dim3(nb), dim3(128), 0, c->stream, sa);
            else hipLaunchKernelGGL((pgfk::k_sens_elem<3>), dim3(nb), dim3(128), 0, c->stream, sa);
            SCHK(hipGetLastError());
        }
        if (a->rhs) {
            const unsigned nb = (unsigned)std::max<int64_t>(1, (c->n_all + 255) / 256);
            if (c->n_cut > 0) {
                if (c->dim == 2)
                    hipLaunchKernelGGL((pgfk::k_sens_rows<2>), dim3(nb), dim3(256), 0, c->stream, c->n_all, c->dof, c->n_free,
                                       c->image, c->inc_ptr, c->inc_elem, c->inc_slot, c->cut_index, c->sens_param,
                                       a->slot, c->sens_pe, drhs);
                else
                    hipLaunchKernelGGL((pgfk::k_sens_rows<3>), dim3(nb), dim3(256), 0, c->stream, c->n_all, c->dof, c->n_free,
                                       c->image, c->inc_ptr, c->inc_elem, c->inc_slot, c->cut_index, c->sens_param,
                                       a->slot, c->sens_pe, drhs);
                SCHK(hipGetLastError());
            } else {
                SCHK(hipMemsetAsync(drhs, 0, sizeof(double) * std::max<int64_t>(c->n_free, 1), c->stream));
            }
            SCHK(hipMemcpyAsync(a->rhs, drhs, sizeof(double) * c->n_free, hipMemcpyDeviceToHost, c->stream));
        }
        std::vector<double> ds((size_t)std::max<int64_t>(c->n_cut, 1) * (S + 1));
        if (a->n_seg > 0 && c->n_cut > 0)
            SCHK(hipMemcpyAsync(ds.data(), c->sens_ds, sizeof(double) * c->n_cut * (S + 1), hipMemcpyDeviceToHost,
                                c->stream));
        SCHK(hipStreamSynchronize(c->stream));
        SCHK(hipFree(deb));
        SCHK(hipFree(drhs));
        if (a->n_seg > 0) {
            // integrals of the explicit stress derivative and the volume
            // derivative per segment, cut elements in order
            std::fill(a->dsig, a->dsig + (size_t)a->n_seg * (S + 1), 0.0);
            for (int64_t ci = 0; ci < c->n_cut; ++ci) {
                if (c->h_sens_param[ci] != a->slot) continue;
                const int64_t e = c->h_cut_elem[ci];
                const int64_t* it = std::upper_bound(a->seg, a->seg + a->n_seg + 1, e);
                const int64_t k = (it - a->seg) - 1;
                if (k < 0 || k >= a->n_seg) continue;
                for (int i = 0; i <= S; ++i) a->dsig[k * (S + 1) + i] += ds[ci * (S + 1) + i];
            }
        }
        return 0;
    }, &args);
}

}  // extern "C"
