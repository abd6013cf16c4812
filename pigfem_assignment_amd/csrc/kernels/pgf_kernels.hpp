// Device kernels of the PIGFEM MI355X path (gfx950, wave64).
//
//  * k_uncut<DIM,...>   one thread per uncut parent element of one colour:
//                       geometry from the nodal coordinates, per-QP B^T D B and
//                       B^T sigma, read-modify-write into the CSR rows (no
//                       atomics: elements of one colour share no node).
//  * k_cut<DIM,...>     one wavefront per cut (IGFEM) element: integration
//                       tables staged in LDS, element matrix accumulated in LDS,
//                       cohesive points (traction-separation law + delta_max
//                       update) folded in, then scattered like k_uncut.
//  * k_spmv_p / k_cg_update   fused Jacobi-PCG iteration (two kernels per
//                       iteration, deterministic last-block reductions).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pgfk {

constexpr int WAVE = 64;
constexpr int MAX_NT = 12;          // element slots of a cut element
constexpr int MAX_M = MAX_NT * 3;   // element dofs
// integration points of one cut element staged in LDS (checked at create)
constexpr int max_vpt(int dim) { return dim == 2 ? 32 : 16; }
constexpr int MAX_CPT = 8;
// threads of the cut-element kernels (one workgroup per cut element)
template <int DIM>
__host__ __device__ constexpr int cut_threads() { return 64; }
// row-gather blocks: threads (slots x matrix rows) and LDS segment (doubles)
constexpr int GATHER_THREADS = 256;
// cut-element scratch of the gather scheme: K_e rows of M entries stored with
// an even stride (16-byte aligned rows, read as pairs), and the rows' column
// bytes with a stride rounded up to 4 (read as 32-bit words)
__host__ __device__ constexpr int cut_ke_stride(int m) { return (m + 1) & ~1; }
__host__ __device__ constexpr int cut_rel_stride(int m) { return (m + 3) & ~3; }
// The scratch holds, per slot of a cut element, its DIM rows of K_e (stride
// cut_ke_stride), the rows' column bytes (cut_rel_stride bytes, placed once at
// set-up) and the slot's DIM entries of P_e: what one incidence of the gather
// reads sits together, addressed by one offset.  Sizes in doubles; a slot
// block has an even size, so every block and row starts 16-byte aligned.
__host__ __device__ constexpr int cut_rel_doubles(int m) { return (cut_rel_stride(m) + 7) / 8; }
__host__ __device__ constexpr int cut_slot_doubles(int m, int dim) {
    return (dim * cut_ke_stride(m) + cut_rel_doubles(m) + dim + 1) & ~1;
}
constexpr int SEG_MAX = 4096;
// MI355X dispatches consecutive workgroups to its 8 XCDs in turn, each XCD
// with its own L2.  Launching 8*per blocks and mapping block b to logical
// block (b % 8) * per + b / 8 gives every XCD one contiguous run of logical
// blocks, so neighbouring rows' element data stay in one L2.
constexpr int N_XCD = 8;
__host__ __device__ inline int64_t xcd_blocks(int64_t n) { return (n + N_XCD - 1) / N_XCD * N_XCD; }
__device__ inline int64_t xcd_logical(int64_t b, int64_t grid) {
    const int64_t per = grid / N_XCD;
    return (b % N_XCD) * per + b / N_XCD;
}

struct CohParams {
    int type;
    double sc, dc, df, beta, visc, tfloor;
};

// Arguments shared by the assembly kernels (passed by value).
struct AsmArgs {
    int64_t n_all, n_free, n_rows, n_uncut, n_cut;
    const double* X;        // [3][n_all]
    const int32_t* dof;     // [dim][n_all]
    const double* U;        // [dim][n_all]
    // uncut elements in colour order
    const int32_t* ue_conn;  // [4][n_uncut]
    const uint8_t* ue_mat;   // [n_uncut]
    const uint32_t* ue_off;  // [8][n_uncut]
    const double* Dmat;      // [n_mat][36]
    double tet_scale;
    // cut elements
    const int32_t* ce_list;      // cut indices in colour order
    const int32_t* ce_nodes;     // [n_cut][16]
    const int32_t* ce_nt;        // [n_cut]
    const int64_t* vpt_ptr;
    const double* vpt_w;
    const int32_t* vpt_mat;
    const int64_t* vpt_grad_off;
    const double* vpt_grad;
    const int64_t* cpt_ptr;
    const double* cpt_w;
    const double* cpt_N;
    const double* cpt_R;
    const uint8_t* cpt_nodes;
    const uint8_t* cpt_nh;
    const int32_t* cpt_law;
    const CohParams* laws;
    const uint8_t* ce_off;
    const int64_t* ce_off_ptr;
    const uint8_t* ce_rel;      // per cut element [nt][nt * dim] row-relative column positions
    const int64_t* ce_rel_ptr;
    const double* dmax_conv;
    double* dmax_trial;
    uint8_t* flags;
    double dt;
    // gather scheme: element-order arrays and the node -> element incidence
    int64_t n_elem;
    const int32_t* el_conn;    // [4][n_elem]
    const uint8_t* el_mat;     // [n_elem]
    const int32_t* cut_index;  // [n_elem], -1 if uncut
    const uint8_t* image;      // [n_all] or null: 1 if the node's dofs belong to an earlier node (periodic image)
    const int64_t* inc_ptr;    // [n_all+1]
    const int32_t* inc_elem;
    const uint8_t* inc_slot;
    const uint8_t* inc_rel;    // [n_inc][4 * dim] row-relative column positions (uncut elements)
    // gather records: per incidence (element, slot | material << 16) or, for a
    // cut element, (the slot's block in the scratch, slot | 0x100 | M << 9 |
    // material << 16); per element its 4 nodes; per node its coordinates
    // padded to 2 or 4 doubles (one or two 16-byte loads)
    const int2* inc_rec;       // [n_inc]
    const int4* el_conn4;      // [n_elem]
    const double* Xq;          // [n_all][2] (2D) or [n_all][4] (3D)
    double* ke;                // cut-element K_e scratch
    const int64_t* ke_ptr;
    double* pe;                // cut-element P_e scratch
    const int64_t* pe_ptr;
    int64_t n_gblk;            // end of the gather blocks of this launch
    int64_t gblk_lo;           // first gather block of this launch
    const int64_t* blk_node;   // [n_gblk+1] first node of each gather block
    const int64_t* blk_seg;    // [n_gblk][2] CSR segment of the block's free rows (empty: not staged)
    const int32_t* blk_rounds; // [n_gblk] incidence rounds: max incident elements of a block node / slots
    // outputs
    const int32_t* rowptr;
    double* K;
    float* K32;
    double* P;            // [n_rows]
    int all_rows;
    int grad_contig;  // 1: a cut element's point gradients are one run starting at vpt_grad_off[first point]
};

// ---- Ortiz-Pandolfi / bilinear traction-separation law ---------------------
// src/Material_OPCohesive.cpp:38-319 (active exponential law with linear
// unloading and viscous term; penetration branch disabled as in the reference,
// :115), delta_eff from src/Material_OPCohesiveNU.cpp:45-63, damage zone from
// :224-236.  Bilinear: the commented law of :175-199/256-271 with delta_c and
// delta_f as parameters.  The irreversibility test and the delta_max update are
// FP64 in every precision mode.  The dimension is a template parameter so
// every small array stays in registers.
template <int dim>
__device__ inline void coh_law(const CohParams& L, double* delta, double dmax, double dt,
                               double* trac, double* Dc, double& dmax_new, uint8_t& flag) {
    const double b2 = L.beta * L.beta;
    double norm = 0.0;
    for (int i = 0; i < dim; ++i) norm += delta[i] * delta[i];
    if (norm < 1e-40)
        for (int i = 0; i < dim; ++i) delta[i] = 1e-20;
    double deff = delta[dim - 1] * delta[dim - 1];
    for (int i = 0; i < dim - 1; ++i) deff += b2 * (delta[i] * delta[i]);
    deff = sqrt(deff);
    double T, dT;
    const bool loading = deff >= dmax || deff < 0.0;
    if (L.type == 0) {
        if (loading) {
            T = L.sc * (deff / L.dc) * exp(1.0 - deff / L.dc);
            dT = (L.sc / L.dc) * (1.0 - deff / L.dc) * exp(1.0 - deff / L.dc);
            if (L.visc != 0.0) {
                T += L.visc / L.sc * (deff - dmax) / dt / L.dc;
                dT += L.visc / L.sc / dt / L.dc;
            }
            dmax_new = deff;
        } else {
            double tmax = L.sc * (dmax / L.dc) * exp(1.0 - dmax / L.dc);
            T = tmax * (deff / dmax);
            dT = tmax / dmax;
            dmax_new = dmax;
        }
    } else {
        auto curve = [&](double x, double& slope) {
            if (x <= L.dc) { slope = L.sc / L.dc; return L.sc / L.dc * x; }
            if (x <= L.df) { slope = -L.sc / (L.df - L.dc); return -L.sc / (L.df - L.dc) * (x - L.dc) + L.sc; }
            slope = 0.0;
            return 0.0;
        };
        if (loading) {
            T = curve(deff, dT);
            if (T < L.tfloor && deff > L.dc) { T = L.tfloor; dT = 0.0; }
            dmax_new = deff;
        } else {
            double s;
            double tmax = curve(dmax, s);
            T = tmax * (deff / dmax);
            dT = tmax / dmax;
            dmax_new = dmax;
        }
    }
    double br[3], dd[3];
    for (int i = 0; i < dim; ++i) br[i] = b2 * delta[i];
    br[dim - 1] += (1.0 - b2) * delta[dim - 1];
    for (int i = 0; i < dim; ++i) trac[i] = (T / deff) * br[i];
    dd[dim - 1] = delta[dim - 1] / deff;
    for (int i = 0; i < dim - 1; ++i) dd[i] = b2 * delta[i] / deff;
    for (int i = 0; i < dim; ++i)
        for (int k = 0; k < dim; ++k) {
            double kr = (i == k ? 1.0 : 0.0);
            double kn = (i == dim - 1 && k == dim - 1) ? 1.0 : 0.0;
            Dc[i * dim + k] = br[i] * (dT * deff - T) * (dd[k] / (deff * deff)) + (T / deff) * (b2 * kr + (1.0 - b2) * kn);
        }
    double nz = deff / L.dc;
    uint8_t zone = nz <= 0 ? 0 : (nz <= 1 ? 1 : (nz <= 5.0 ? 2 : 3));
    flag = (uint8_t)((loading ? 1 : 0) | (zone << 1) | (dmax_new > L.dc ? 8 : 0) |
                     ((L.type == 1 && dmax_new >= L.df) ? 16 : 0));
}

}  // namespace pgfk
