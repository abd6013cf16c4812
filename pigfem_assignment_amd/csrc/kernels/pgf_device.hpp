// Device-resident context of one mesh (what pgf_ctx points to).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "pgf_kernels.hpp"

struct pgf_ctx;  // the C-ABI handle (derives from pgfd::Ctx)

namespace pgfd {

struct LearnState;  // learned initial guess state (pgf_cnn.hip)

struct Ctx {
    std::string err;
    hipStream_t stream = nullptr;
    int dim = 0, kind = 0, n_colors = 0;
    double tet_scale = 1.0;
    int64_t n_nodes = 0, n_all = 0, n_elem = 0, n_uncut = 0, n_cut = 0, n_free = 0, n_pres = 0, n_rows = 0,
            nnz = 0, nnz_ff = 0, n_cpt = 0, n_mat = 0;
    // host copies of the pattern (exported for parity checks)
    std::vector<int64_t> h_rowptr;
    std::vector<int32_t> h_rowsplit, h_colidx, h_color;
    std::vector<int64_t> ue_color_ptr, ce_color_ptr;
    // mesh
    double* X = nullptr;
    int32_t* dof = nullptr;
    double* U = nullptr;
    int32_t* ue_conn = nullptr;
    uint8_t* ue_mat = nullptr;
    uint32_t* ue_off = nullptr;
    double* Dmat = nullptr;
    int32_t *ce_list = nullptr, *ce_nodes = nullptr, *ce_nt = nullptr;
    int64_t *vpt_ptr = nullptr, *vpt_grad_off = nullptr, *cpt_ptr = nullptr, *ce_off_ptr = nullptr;
    double *vpt_w = nullptr, *vpt_grad = nullptr, *cpt_w = nullptr, *cpt_N = nullptr, *cpt_R = nullptr;
    int32_t *vpt_mat = nullptr, *cpt_law = nullptr;
    uint8_t *cpt_nodes = nullptr, *cpt_nh = nullptr, *ce_off = nullptr;
    pgfk::CohParams* laws = nullptr;
    double *dmax_conv = nullptr, *dmax_trial = nullptr;
    uint8_t* flags = nullptr;
    // gather scheme
    int32_t *el_conn = nullptr, *cut_index = nullptr, *inc_elem = nullptr;
    int32_t *inc_rec = nullptr, *el_conn4 = nullptr;  // gather records (see AsmArgs)
    double* Xq = nullptr;
    uint8_t *el_mat = nullptr, *inc_slot = nullptr;
    int64_t *inc_ptr = nullptr, *ke_ptr = nullptr, *pe_ptr = nullptr;
    uint8_t* inc_rel = nullptr;  // [n_inc][4 * dim] row-relative column positions (uncut elements)
    uint8_t* ce_rel = nullptr;   // per cut element [nt][nt * dim] row-relative column positions
    int64_t* ce_rel_ptr = nullptr;
    double *ke = nullptr, *pe = nullptr;
    int64_t n_inc = 0, n_ke = 0;
    int64_t *blk_node = nullptr, *blk_seg = nullptr;
    int32_t* blk_rounds = nullptr;
    int64_t n_gblk = 0, seg_max = 0;
    int64_t n_gblk_reg = 0;         // gather blocks of the regular nodes (the rest: enrichment nodes)
    int gslots = 1;                 // gather pass: threads per row (or per node with gnode)
    bool gnode = false;             // gather pass: one thread serves all rows of a node
    bool grad_contig = false;       // every cut element's gradient blocks are consecutive
    int64_t n_vpt = 0, n_vgrad = 0;  // volumetric points of the cut elements and their gradient entries
    std::vector<int32_t> h_cut_elem; // [n_cut] parent element of each cut element
    // shape-sensitivity tables and scratch (pgf_sens.hip)
    double *sens_dgrad = nullptr, *sens_divv = nullptr, *sens_pe = nullptr, *sens_ds = nullptr;
    int32_t* sens_param = nullptr;
    std::vector<int32_t> h_sens_param;
    uint8_t* image = nullptr;        // [n_all] periodic images (null without shared dofs)
    std::vector<uint8_t> h_image;    // host copy (empty without shared dofs)
    std::vector<int32_t> h_dof;      // [dim][n_all] (gather partition)
    std::vector<int64_t> h_inc_ptr;  // [n_all+1] node -> element incidence
    // free dof -> nodal field slot (d*n_all + node) of the first node holding it
    int64_t* free_slot = nullptr;
    // matrix
    int32_t *rowptr = nullptr, *rowsplit = nullptr, *colidx = nullptr;
    int32_t* grp_row = nullptr;  // [n_grp+1] first free row of each SpMV node group
    int64_t n_grp = 0;
    double* K = nullptr;
    float* K32 = nullptr;
    double* P = nullptr;
    double* Fext_full = nullptr;   // [n_free]
    double* upres_full = nullptr;  // [n_pres]
    // solver work (FP64 and FP32)
    double *r = nullptr, *x = nullptr, *z = nullptr, *p0 = nullptr, *p1 = nullptr, *q = nullptr, *dinv = nullptr,
           *guess = nullptr, *res = nullptr;
    double* partial = nullptr;     // block partial sums
    double* scal = nullptr;        // device scalars of the PCG recurrence
    unsigned* counter = nullptr;   // last-block counters
    int* done = nullptr;
    double* h_scal = nullptr;      // pinned mirror
    bool have_guess = false;
    hipEvent_t* timing_events = nullptr;  // when set: cut pass start, gather start, gather end
    LearnState* learn = nullptr;

    pgfk::AsmArgs asm_args() const;
    ~Ctx();
};

template <typename T>
void fill_async(Ctx& c, T* p, int64_t n, T v);
void enqueue_assembly(Ctx& c, double dt, uint32_t flags);
// homogenised response per element segment (partials [n_seg * 64][2S+1], out [n_seg][2S+1])
void enqueue_volume_average(Ctx& c, const double* d_eps_bar, int n_seg, const int64_t* d_seg, double* d_part,
                            double* d_out);
// the current device's scratch of the allocation-free entry point
double* noalloc_scratch();
// nodal field U from a dof vector (free part x, prescribed part up or zero)
void enqueue_set_from_dofs(Ctx& c, const double* d_x, const double* d_up);
int default_gather_slots(int dim);
void build_gather_blocks(Ctx& c, int slots, bool node_rows);

// Learned initial guess (pgf_cnn.hip).  Called after the Jacobian assembly of
// Newton iteration `it` with that iteration's residual in c.r.  Iteration 0
// records the step-start features and U (training samples).  For use_guess 2
// (network + previous increment / previous Newton correction) or 3 (previous
// increment / correction only) it writes the Galerkin-scaled guess of this
// iteration's solve to c.guess and returns true.
bool learn_guess(Ctx& c, int use_guess, bool fp32, int it);
void learn_free(LearnState* s);
Ctx* ctx_of(pgf_ctx* h);

}  // namespace pgfd
