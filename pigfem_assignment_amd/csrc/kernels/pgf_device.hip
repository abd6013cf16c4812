// Device context, Krylov kernels, Newton step and the pgf_* C-ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>

#include "../../../include/pgfem.h"
#include "../host/pgf_dofs.hpp"
#include "pgf_device.hpp"

namespace pgfd {

#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            throw pgf::Error(pgf::PGF_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

static std::string g_last_error;

// Device scratch of the allocation-free entry point (pgf_volume_average_noalloc):
// staging of eps_bar and the segment list, the per-segment results and the
// per-block partial sums, sized for PGF_NOALLOC_MAX_SEG segments.  Every GPU
// holds its own copy; host code uses the address hipGetSymbolAddress returns
// for the current device, never the array's name.
constexpr size_t NOALLOC_DOUBLES =
    8 + (PGF_NOALLOC_MAX_SEG + 1) + (size_t)PGF_NOALLOC_MAX_SEG * 13 + (size_t)PGF_NOALLOC_MAX_SEG * 64 * 13;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
static std::mutex g_scratch_mutex;
static double* g_scratch_ptr[64] = {nullptr};

double* noalloc_scratch() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) throw pgf::Error(PGF_EDEVICE, "no current device");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw pgf::Error(PGF_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (double*)ptr;
    }
    return g_scratch_ptr[dev];
}

template <typename T>
static T* dalloc(int64_t n) {
    T* p = nullptr;
    if (n <= 0) n = 1;
    HIPCHK(hipMalloc((void**)&p, sizeof(T) * (size_t)n));
    return p;
}
template <typename T>
static T* dupload(const T* h, int64_t n) {
    T* p = dalloc<T>(n);
    if (n > 0 && h) HIPCHK(hipMemcpy(p, h, sizeof(T) * (size_t)n, hipMemcpyHostToDevice));
    return p;
}

template <typename T>
__global__ void k_fill(T* p, int64_t n, T v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
template <typename T>
void fill_async(Ctx& c, T* p, int64_t n, T v) {
    if (n <= 0) return;
    int64_t nb = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)nb), dim3(256), 0, c.stream, p, n, v);
}
template void fill_async<double>(Ctx&, double*, int64_t, double);
template void fill_async<float>(Ctx&, float*, int64_t, float);

pgfk::AsmArgs Ctx::asm_args() const {
    pgfk::AsmArgs a{};
    a.n_all = n_all; a.n_free = n_free; a.n_rows = n_rows; a.n_uncut = n_uncut;
    a.X = X; a.dof = dof; a.U = U;
    a.ue_conn = ue_conn; a.ue_mat = ue_mat; a.ue_off = ue_off; a.Dmat = Dmat; a.tet_scale = tet_scale;
    a.ce_list = ce_list; a.ce_nodes = ce_nodes; a.ce_nt = ce_nt;
    a.vpt_ptr = vpt_ptr; a.vpt_w = vpt_w; a.vpt_mat = vpt_mat; a.vpt_grad_off = vpt_grad_off; a.vpt_grad = vpt_grad;
    a.cpt_ptr = cpt_ptr; a.cpt_w = cpt_w; a.cpt_N = cpt_N; a.cpt_R = cpt_R; a.cpt_nodes = cpt_nodes; a.cpt_nh = cpt_nh;
    a.cpt_law = cpt_law; a.laws = laws; a.ce_off = ce_off; a.ce_off_ptr = ce_off_ptr; a.ce_rel = ce_rel; a.ce_rel_ptr = ce_rel_ptr;
    a.dmax_conv = dmax_conv; a.dmax_trial = dmax_trial; a.flags = flags;
    a.rowptr = rowptr; a.K = K; a.K32 = K32; a.P = P;
    a.n_elem = n_elem; a.el_conn = el_conn; a.el_mat = el_mat; a.cut_index = cut_index; a.image = image;
    a.inc_ptr = inc_ptr; a.inc_elem = inc_elem; a.inc_slot = inc_slot; a.inc_rel = inc_rel;
    a.inc_rec = (const int2*)inc_rec; a.el_conn4 = (const int4*)el_conn4; a.Xq = Xq;
    a.grad_contig = grad_contig ? 1 : 0;
    a.ke = ke; a.ke_ptr = ke_ptr; a.pe = pe; a.pe_ptr = pe_ptr;
    a.n_gblk = n_gblk; a.blk_node = blk_node; a.blk_seg = blk_seg; a.blk_rounds = blk_rounds; a.n_cut = n_cut;
    return a;
}

Ctx::~Ctx() {
    void* ptrs[] = {X, dof, U, ue_conn, ue_mat, ue_off, Dmat, ce_list, ce_nodes, ce_nt, vpt_ptr, vpt_grad_off, cpt_ptr,
                    ce_off_ptr, vpt_w, vpt_grad, cpt_w, cpt_N, cpt_R, vpt_mat, cpt_law, cpt_nodes, cpt_nh, ce_off, laws,
                    dmax_conv, dmax_trial, flags, free_slot, rowptr, rowsplit, colidx, K, K32, P, Fext_full,
                    upres_full, r, x, z, p0, p1, q, dinv, guess, res, partial,
                    scal, counter, done, el_conn, cut_index, inc_elem, el_mat, inc_slot, inc_ptr, ke_ptr, pe_ptr,
                    inc_rel, ke, pe, blk_node, blk_seg, blk_rounds, grp_row, ce_rel, ce_rel_ptr, image,
                    sens_dgrad, sens_divv, sens_pe, sens_ds, sens_param, inc_rec, el_conn4, Xq};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (h_scal) (void)hipHostFree(h_scal);
    if (learn) learn_free(learn);
    if (stream) (void)hipStreamDestroy(stream);
}

// ---------------------------------------------------------------------------
// reductions: wave64 shuffles, then LDS across the block's waves
template <int NV>
__device__ inline void block_reduce(double* v) {
    __shared__ double sh[NV][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    for (int k = 0; k < NV; ++k)
        for (int m = 32; m >= 1; m >>= 1) v[k] += __shfl_xor(v[k], m, 64);
    __syncthreads();
    if (lane == 0)
        for (int k = 0; k < NV; ++k) sh[k][wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += sh[k][w];
            v[k] = s;
        }
}

// Publish a block's partial sums; returns true in the one block that arrives
// last, after which it may read every block's partials (fixed-order sum).
template <int NV>
__device__ inline bool last_block(double* v, double* partial, unsigned* counter) {
    __shared__ bool s_last;
    block_reduce<NV>(v);
    if (threadIdx.x == 0) {
        for (int k = 0; k < NV; ++k) partial[k * gridDim.x + blockIdx.x] = v[k];
        __threadfence();
        unsigned t = atomicAdd(counter, 1u);
        s_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!s_last) return false;
    __threadfence();
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x) s += partial[k * gridDim.x + i];
        v[k] = s;
    }
    block_reduce<NV>(v);
    return true;
}

// One block's sums to partial[k * gridDim.x + blockIdx.x]; the next kernel on
// the stream reads them.
template <int NV>
__device__ inline void block_partial(double* v, double* partial) {
    block_reduce<NV>(v);
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; ++k) partial[k * gridDim.x + blockIdx.x] = v[k];
}

// Fixed-order sum of nb partial sums of NV values by one block; four
// independent accumulators per thread keep four loads in flight.
template <int NV>
__device__ inline void sum_partials(int nb, const double* partial, double* v) {
    const int bd = blockDim.x;
    for (int k = 0; k < NV; ++k) {
        const double* pk = partial + (int64_t)k * nb;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        int i = threadIdx.x;
        for (; i + 3 * bd < nb; i += 4 * bd) {
            s0 += pk[i];
            s1 += pk[i + bd];
            s2 += pk[i + 2 * bd];
            s3 += pk[i + 3 * bd];
        }
        for (; i < nb; i += bd) s0 += pk[i];
        v[k] = (s0 + s1) + (s2 + s3);
    }
    block_reduce<NV>(v);
}

// Gather blocks: runs of consecutive nodes, `slots` threads per matrix row
// (node_rows = 0) or per node (node_rows = 1, one thread serves all rows of a
// node), GATHER_THREADS / (slots * dim) or GATHER_THREADS / slots nodes per
// block.  A block's free rows are
// staged in LDS when they form one contiguous CSR segment of at most SEG_MAX
// entries; blk_rounds holds the incident elements of the block's busiest node
// divided by the slots (rounded up).
int default_gather_slots(int dim) { return dim == 2 ? 1 : 2; }

void build_gather_blocks(Ctx& c, int slots, bool node_rows) {
    if (slots != 1 && slots != 2 && slots != 4 && slots != 8)
        throw pgf::Error(PGF_EARG, "gather slots must be 1, 2, 4 or 8");
    const int per_blk = pgfk::GATHER_THREADS / (slots * (node_rows ? 1 : c.dim));
    const int64_t na = c.n_all;
    const std::vector<int64_t>& rp = c.h_rowptr;
    std::vector<int64_t> bn{0}, bs;
    std::vector<int32_t> brounds;
    int64_t seg_max = 0;
    for (int64_t n0 = 0; n0 < na;) {
        int64_t n1 = n0, rmin = INT64_MAX, rmax = -1, cnt = 0, inc_max = 0;
        // blocks do not straddle the last regular node: regular and enrichment
        // nodes are gathered by separate launches
        while (n1 < na && n1 - n0 < per_blk && !(n1 == c.n_nodes && n1 > n0)) {
            int64_t mn = rmin, mx = rmax, ct = cnt;
            // a periodic image owns no rows: the node holding its dofs first does
            const bool owns = c.h_image.empty() || !c.h_image[n1];
            for (int dd = 0; dd < c.dim && owns; ++dd) {
                const int64_t g = c.h_dof[(size_t)dd * na + n1];
                if (g < c.n_free) { mn = std::min(mn, g); mx = std::max(mx, g); ++ct; }
            }
            if (n1 > n0 && ct > 0 && rp[mx + 1] - rp[mn] > pgfk::SEG_MAX) break;
            rmin = mn; rmax = mx; cnt = ct;
            inc_max = std::max<int64_t>(inc_max, c.h_inc_ptr[n1 + 1] - c.h_inc_ptr[n1]);
            ++n1;
        }
        brounds.push_back((int32_t)((inc_max + slots - 1) / slots));
        int64_t s0 = 0, s1 = 0;
        if (cnt > 0 && rmax - rmin + 1 == cnt && rp[rmax + 1] - rp[rmin] <= pgfk::SEG_MAX) {
            s0 = rp[rmin];
            s1 = rp[rmax + 1];
            seg_max = std::max(seg_max, s1 - s0);
        }
        bs.push_back(s0);
        bs.push_back(s1);
        bn.push_back(n1);
        n0 = n1;
    }
    int64_t n_reg_blk = 0;
    while (n_reg_blk + 1 < (int64_t)bn.size() && bn[n_reg_blk + 1] <= c.n_nodes) ++n_reg_blk;
    if ((int64_t)bn.size() - 1 > 0x7fffffff) throw pgf::Error(PGF_ELIMIT, "too many gather blocks");
    int64_t* node = dupload(bn.data(), (int64_t)bn.size());
    int64_t* seg = dupload(bs.data(), std::max<int64_t>((int64_t)bs.size(), 2));
    int32_t* rounds = dupload(brounds.data(), std::max<int64_t>((int64_t)brounds.size(), 1));
    if (c.blk_node) HIPCHK(hipFree(c.blk_node));
    if (c.blk_seg) HIPCHK(hipFree(c.blk_seg));
    if (c.blk_rounds) HIPCHK(hipFree(c.blk_rounds));
    c.blk_node = node; c.blk_seg = seg; c.blk_rounds = rounds;
    c.n_gblk = (int64_t)bn.size() - 1;
    c.n_gblk_reg = n_reg_blk;
    c.seg_max = seg_max;
    c.gslots = slots;
    c.gnode = node_rows;
}

// scalar slots of the PCG recurrence
enum { S_RZ = 0, S_ALPHA, S_BETA, S_RR, S_RR0, S_TOL2, S_ITER, S_MAXIT, S_BB, S_NSCAL = 16 };

constexpr int RED_BLOCKS = 1024;
// PCG iteration kernels write one partial sum per block and a one-block kernel
// sums them in a fixed order: no atomics or fences in the iteration.  A single
// completion counter per kernel costs ~60 ns per block on one L2 address,
// which at 2,000-30,000 blocks was most of an iteration.
constexpr int CG_SPMV_BLOCKS = 65536;   // cap of the SpMV grid (partials slots)
constexpr int CG_UPD_BLOCKS = 2048;     // grid of the vector update
constexpr int CG_PART = CG_SPMV_BLOCKS + 2 * CG_UPD_BLOCKS;  // partial slots of one context

// ---------------------------------------------------------------------------
// SpMV on the free-free block, y = K_ff x, by node groups: the free rows of
// one node are consecutive and have the same free column list, so G lanes
// serve a group of up to three rows together and every column index and x
// entry is loaded once for all of them (the column index stream shrinks to a
// third in 3D).  Each row still sums its entries lane-strided and then across
// the G lanes, the same order as one row per G lanes.
// Places each cut element's column bytes (compact rows of cut_rel_stride
// bytes) into its slots' blocks of the gather scratch, once at set-up: the
// cut kernel writes only the K_e rows and P_e entries of a block.
__global__ __launch_bounds__(256) void k_place_rel(int64_t n_cut, int dim, const int32_t* __restrict__ ce_nt,
                                                   const uint8_t* __restrict__ rel, const int64_t* __restrict__ rel_ptr,
                                                   const int64_t* __restrict__ ke_ptr, double* __restrict__ ke) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n_cut) return;
    const int nt = ce_nt[k], m = nt * dim, rs = pgfk::cut_rel_stride(m);
    const int sd = pgfk::cut_slot_doubles(m, dim), ks = pgfk::cut_ke_stride(m);
    for (int la = 0; la < nt; ++la) {
        const uint32_t* src = (const uint32_t*)(rel + rel_ptr[k] + (int64_t)la * rs);
        uint32_t* dst = (uint32_t*)(ke + ke_ptr[k] + (int64_t)la * sd + (int64_t)dim * ks);
        for (int w = 0; w < rs / 4; ++w) dst[w] = src[w];
    }
}

template <typename TA, typename T, int G>
__global__ __launch_bounds__(256) void k_spmv(int64_t ng, const int32_t* __restrict__ grow,
                                              const int32_t* __restrict__ rowptr, const int32_t* __restrict__ rowsplit,
                                              const int32_t* __restrict__ col, const TA* __restrict__ A,
                                              const T* __restrict__ x, T* __restrict__ y) {
    const int lane = threadIdx.x % G;
    for (int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; g < ng;
         g += (int64_t)gridDim.x * blockDim.x / G) {
        const int32_t r0 = grow[g], k = grow[g + 1] - r0;
        const int32_t b0 = rowptr[r0], len = rowsplit[r0];
        const int32_t b1 = k > 1 ? rowptr[r0 + 1] : b0, b2 = k > 2 ? rowptr[r0 + 2] : b0;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int32_t j = lane; j < len; j += G) {
            const double xv = (double)x[col[b0 + j]];
            a0 += (double)A[b0 + j] * xv;
            if (k > 1) a1 += (double)A[b1 + j] * xv;
            if (k > 2) a2 += (double)A[b2 + j] * xv;
        }
        for (int m = G / 2; m >= 1; m >>= 1) {
            a0 += __shfl_xor(a0, m, G);
            a1 += __shfl_xor(a1, m, G);
            a2 += __shfl_xor(a2, m, G);
        }
        if (lane == 0) {
            y[r0] = (T)a0;
            if (k > 1) y[r0 + 1] = (T)a1;
            if (k > 2) y[r0 + 2] = (T)a2;
        }
    }
}

// PCG step A1: p = z + beta p (elementwise, in place).
template <typename T>
__global__ __launch_bounds__(256) void k_cg_pupd(int64_t n, const T* __restrict__ z, T* __restrict__ p,
                                                 const double* scal, const int* done) {
    if (*done) return;
    const double beta = scal[S_BETA];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (T)((double)z[i] + beta * (double)p[i]);
}

// PCG step A2: q = A p by node groups (each column index and p entry loaded
// once per group), per-block partial of p.q.  The matrix may be stored in a
// narrower type than the Krylov vectors.
template <typename TA, typename T, int G>
__global__ __launch_bounds__(256) void k_cg_spmv(int64_t ng, const int32_t* __restrict__ grow,
                                                 const int32_t* __restrict__ rowptr,
                                                 const int32_t* __restrict__ rowsplit, const int32_t* __restrict__ col,
                                                 const TA* __restrict__ A, const T* __restrict__ p, T* __restrict__ q,
                                                 double* partial, const int* done) {
    if (*done) return;
    const int lane = threadIdx.x % G;
    double dot = 0.0;
    for (int64_t g = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G; g < ng;
         g += (int64_t)gridDim.x * blockDim.x / G) {
        const int32_t r0 = grow[g], k = grow[g + 1] - r0;
        const int32_t b0 = rowptr[r0], len = rowsplit[r0];
        const int32_t b1 = k > 1 ? rowptr[r0 + 1] : b0, b2 = k > 2 ? rowptr[r0 + 2] : b0;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int32_t j = lane; j < len; j += G) {
            const double pv = (double)p[col[b0 + j]];
            a0 += (double)A[b0 + j] * pv;
            if (k > 1) a1 += (double)A[b1 + j] * pv;
            if (k > 2) a2 += (double)A[b2 + j] * pv;
        }
        for (int m = G / 2; m >= 1; m >>= 1) {
            a0 += __shfl_xor(a0, m, G);
            a1 += __shfl_xor(a1, m, G);
            a2 += __shfl_xor(a2, m, G);
        }
        if (lane == 0) {
            const double av[3] = {a0, a1, a2};
            for (int t = 0; t < k; ++t) {
                q[r0 + t] = (T)av[t];
                dot += (double)p[r0 + t] * av[t];
            }
        }
    }
    double v[1] = {dot};
    block_partial<1>(v, partial);
}

// alpha = rz / (p.q) from the SpMV blocks' partials.
__global__ __launch_bounds__(1024) void k_cg_alpha(int nb, const double* __restrict__ partial, double* scal,
                                                   const int* done) {
    if (*done) return;
    double v[1];
    sum_partials<1>(nb, partial, v);
    if (threadIdx.x == 0) {
        const double pq = v[0];
        scal[S_ALPHA] = (pq != 0.0 && std::isfinite(pq)) ? scal[S_RZ] / pq : 0.0;
    }
}

// PCG step B: x += alpha p, r -= alpha q, z = D^{-1} r, then beta and the
// convergence flag in the last block.
template <typename T>
__global__ __launch_bounds__(256) void k_cg_update(int64_t n, T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                   const T* __restrict__ p, const T* __restrict__ q,
                                                   const T* __restrict__ dinv, const double* scal, double* partial,
                                                   const int* done) {
    if (*done) return;
    const double alpha = scal[S_ALPHA];
    double rz = 0.0, rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double xi = (double)x[i] + alpha * (double)p[i];
        const double ri = (double)r[i] - alpha * (double)q[i];
        const double zi = (double)dinv[i] * ri;
        x[i] = (T)xi; r[i] = (T)ri; z[i] = (T)zi;
        rz += ri * zi;
        rr += ri * ri;
    }
    double v[2] = {rz, rr};
    block_partial<2>(v, partial);
}

// beta = rz_new / rz_old, the iteration count and the convergence flag from
// the update blocks' partials.
__global__ __launch_bounds__(1024) void k_cg_beta(int nb, const double* __restrict__ partial, double* scal, int* done) {
    if (*done) return;
    double v[2];
    sum_partials<2>(nb, partial, v);
    if (threadIdx.x == 0) {
        const double rz_old = scal[S_RZ];
        scal[S_BETA] = rz_old != 0.0 ? v[0] / rz_old : 0.0;
        scal[S_RZ] = v[0];
        scal[S_RR] = v[1];
        const double it = scal[S_ITER] + 1.0;
        scal[S_ITER] = it;
        if (!(v[1] > scal[S_TOL2] * scal[S_RR0]) || it >= scal[S_MAXIT] || !std::isfinite(v[1])) *done = 1;
    }
}

// PCG start: r = b - Ax (Ax precomputed in q, or zero), z = D^{-1} r, p = 0.
template <typename T, typename TB>
__global__ __launch_bounds__(256) void k_cg_init(int64_t n, const TB* __restrict__ b, T* __restrict__ r,
                                                 T* __restrict__ z, T* __restrict__ p0, T* __restrict__ p1,
                                                 const T* __restrict__ Ax, const T* __restrict__ dinv, double* scal,
                                                 double* partial, unsigned* counter, int* done, double tol2,
                                                 double maxit) {
    double rz = 0.0, rr = 0.0, bb = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double bi = (double)b[i];
        const double ri = Ax ? bi - (double)Ax[i] : bi;
        const double zi = (double)dinv[i] * ri;
        r[i] = (T)ri; z[i] = (T)zi; p0[i] = (T)0; p1[i] = (T)0;
        rz += ri * zi; rr += ri * ri; bb += bi * bi;
    }
    double v[3] = {rz, rr, bb};
    if (last_block<3>(v, partial, counter + 2) && threadIdx.x == 0) {
        scal[S_RZ] = v[0]; scal[S_RR] = v[1]; scal[S_BB] = v[2];
        // relative to ||b|| (KSP's default preconditioned/unpreconditioned choice is ||b||)
        scal[S_RR0] = v[2];
        scal[S_TOL2] = tol2; scal[S_BETA] = 0.0; scal[S_ALPHA] = 0.0; scal[S_ITER] = 0.0; scal[S_MAXIT] = maxit;
        *done = (!(v[1] > tol2 * v[2]) || maxit <= 0) ? 1 : 0;
        counter[2] = 0;
    }
}

template <typename TA, typename T>
__global__ void k_diag_inv(int64_t n, const int32_t* rowptr, const int32_t* rowsplit, const int32_t* col,
                           const TA* A, T* dinv) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = rowptr[r], e = b + rowsplit[r];
        double d = 0.0;
        for (int32_t j = b; j < e; ++j)
            if (col[j] == r) { d = (double)A[j]; break; }
        dinv[r] = (T)(d != 0.0 ? 1.0 / d : 1.0);
    }
}

// ---- Newton helpers ----
// prescribed values into the nodal field: every node slot (d * n_all + node)
// whose dof is prescribed, so all nodes sharing a dof (periodic images) get it
__global__ void k_apply_frac(int64_t n_slots, const int32_t* dof, int64_t n_free, const double* upf, double frac,
                             double* U) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_slots; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t g = dof[i];
        if (g >= n_free) U[i] = frac * upf[g - n_free];
    }
}
// r = frac*Fext - P_int (free rows) and ||r||^2 into scal[slot]
__global__ __launch_bounds__(256) void k_residual(int64_t n, const double* Fext, double frac, const double* P,
                                                  double* r, double* partial, unsigned* counter, double* out) {
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double ri = frac * Fext[i] - P[i];
        r[i] = ri;
        s += ri * ri;
    }
    double v[1] = {s};
    if (last_block<1>(v, partial, counter) && threadIdx.x == 0) {
        *out = v[0];
        *counter = 0;
    }
}
// U += dx over every node slot with a free dof (periodic images included)
template <typename T>
__global__ void k_update_U(int64_t n_slots, const int32_t* dof, int64_t n_free, const T* dx, double* U) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_slots; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t g = dof[i];
        if (g < n_free) U[i] += (double)dx[g];
    }
}
__global__ void k_copy(int64_t n, const double* a, double* b) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) b[i] = a[i];
}

static unsigned grid_for(int64_t n, int per_block = 256) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + per_block - 1) / per_block, RED_BLOCKS));
}

template <typename TA, typename T>
static void enqueue_spmv(Ctx& c, const TA* A, const T* x, T* y) {
    const unsigned nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>((c.n_grp * 16 + 255) / 256, 65536));
    if (c.dim == 3)
        hipLaunchKernelGGL((k_spmv<TA, T, 16>), dim3(nb), dim3(256), 0, c.stream, c.n_grp, c.grp_row, c.rowptr, c.rowsplit,
                           c.colidx, A, x, y);
    else
        hipLaunchKernelGGL((k_spmv<TA, T, 8>), dim3(nb), dim3(256), 0, c.stream, c.n_grp, c.grp_row, c.rowptr, c.rowsplit,
                           c.colidx, A, x, y);
}

// Enqueue a whole PCG solve in chunks; the device flag stops the recurrence,
// the host only polls the pinned scalars between chunks.
static double wall_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename TA, typename T>
static int run_pcg(Ctx& c, const TA* A, const void* b, bool b_is_double, T* x, T* r, T* z, T* p0, T* p1, T* q,
                   T* dinv, bool x0, double rtol, int maxit, pgf_solve_info* info) {
    const int64_t n = c.n_free;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, c.stream));
    hipLaunchKernelGGL((k_diag_inv<TA, T>), dim3(grid_for(n)), dim3(256), 0, c.stream, n, c.rowptr, c.rowsplit, c.colidx, A, dinv);
    if (x0) enqueue_spmv<TA, T>(c, A, x, q);
    else fill_async<T>(c, x, n, (T)0);
    const unsigned gi = grid_for(n);
    if (b_is_double)
        hipLaunchKernelGGL((k_cg_init<T, double>), dim3(gi), dim3(256), 0, c.stream, n, (const double*)b, r, z, p0, p1,
                           x0 ? (const T*)q : nullptr, (const T*)dinv, c.scal, c.partial, c.counter, c.done, rtol * rtol,
                           (double)maxit);
    else
        hipLaunchKernelGGL((k_cg_init<T, float>), dim3(gi), dim3(256), 0, c.stream, n, (const float*)b, r, z, p0, p1,
                           x0 ? (const T*)q : nullptr, (const T*)dinv, c.scal, c.partial, c.counter, c.done, rtol * rtol,
                           (double)maxit);
    const int G = c.dim == 3 ? 16 : 8;
    const unsigned ga = (unsigned)std::max<int64_t>(1, std::min<int64_t>((c.n_grp * G + 255) / 256, CG_SPMV_BLOCKS));
    const unsigned gu = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, CG_UPD_BLOCKS));
    double* part_a = c.partial;                   // [ga] p.q
    double* part_u = c.partial + CG_SPMV_BLOCKS;  // [2][gu] rz, rr
    const int chunk = 32;
    int it = 0;
    bool stop = false;
    const char* pe = std::getenv("PGF_PROGRESS");
    const double progress_s = pe ? std::atof(pe) : 0.0;
    double last_report = wall_s();
    while (!stop) {
        for (int k = 0; k < chunk; ++k, ++it) {
            T* pnew = p0;  // p is updated in place
            hipLaunchKernelGGL((k_cg_pupd<T>), dim3(gu), dim3(256), 0, c.stream, n, (const T*)z, pnew,
                               (const double*)c.scal, (const int*)c.done);
            if (G == 16)
                hipLaunchKernelGGL((k_cg_spmv<TA, T, 16>), dim3(ga), dim3(256), 0, c.stream, c.n_grp, c.grp_row, c.rowptr,
                                   c.rowsplit, c.colidx, A, (const T*)pnew, q, part_a, (const int*)c.done);
            else
                hipLaunchKernelGGL((k_cg_spmv<TA, T, 8>), dim3(ga), dim3(256), 0, c.stream, c.n_grp, c.grp_row, c.rowptr,
                                   c.rowsplit, c.colidx, A, (const T*)pnew, q, part_a, (const int*)c.done);
            hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(1024), 0, c.stream, (int)ga, (const double*)part_a, c.scal,
                               (const int*)c.done);
            hipLaunchKernelGGL((k_cg_update<T>), dim3(gu), dim3(256), 0, c.stream, n, x, r, z, pnew, q, dinv,
                               (const double*)c.scal, part_u, (const int*)c.done);
            hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(1024), 0, c.stream, (int)gu, (const double*)part_u, c.scal, c.done);
        }
        HIPCHK(hipMemcpyAsync(c.h_scal, c.scal, sizeof(double) * S_NSCAL, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipMemcpyAsync(c.h_scal + S_NSCAL, c.done, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        int dn;
        std::memcpy(&dn, c.h_scal + S_NSCAL, sizeof(int));
        if (dn || it >= maxit) stop = true;
        // optional progress report of long solves (PGF_PROGRESS = seconds between lines)
        if (progress_s > 0.0) {
            const double now = wall_s();
            if (now - last_report >= progress_s) {
                last_report = now;
                std::fprintf(stderr, "[pcg] %d iterations, ||r|| / ||r0|| = %.3e\n", (int)c.h_scal[S_ITER],
                             c.h_scal[S_RR0] > 0 ? std::sqrt(c.h_scal[S_RR] / c.h_scal[S_RR0]) : 0.0);
                std::fflush(stderr);
            }
        }
    }
    HIPCHK(hipEventRecord(e1, c.stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double rr = c.h_scal[S_RR], rr0 = c.h_scal[S_RR0];
    if (info) {
        info->iterations = (int)c.h_scal[S_ITER];
        info->rel_residual = rr0 > 0 ? std::sqrt(rr / rr0) : 0.0;
        info->converged = rr0 == 0 || rr <= rtol * rtol * rr0;
        info->ms = ms;
    }
    // the last iterate's p lives in the buffer written by the final step A
    return 0;
}

}  // namespace pgfd

using pgfd::Ctx;

// ===========================================================================
// C-ABI
// ===========================================================================
namespace {
int fail_ctx(Ctx* c, const pgf::Error& e) {
    if (c) c->err = e.what();
    pgfd::g_last_error = e.what();
    return e.code;
}
int fail_ctx(Ctx* c, const std::exception& e) {
    if (c) c->err = e.what();
    pgfd::g_last_error = e.what();
    return PGF_EARG;
}
template <typename F>
int guard(Ctx* c, F&& f) {
    try {
        return f();
    } catch (const pgf::Error& e) {
        return fail_ctx(c, e);
    } catch (const std::exception& e) {
        return fail_ctx(c, e);
    }
}
}  // namespace

extern "C" {

struct pgf_ctx : public Ctx {};

}  // extern "C"

pgfd::Ctx* pgfd::ctx_of(pgf_ctx* h) { return h; }

extern "C" {

PGF_API const char* pgf_last_error(const pgf_ctx* ctx) {
    return ctx ? ctx->err.c_str() : pgfd::g_last_error.c_str();
}

PGF_API int pgf_device_info(char* name, int32_t name_len, int32_t* n_cu, double* hbm_gb) {
    return guard(nullptr, [&]() -> int {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        hipDeviceProp_t p;
        HIPCHK(hipGetDeviceProperties(&p, dev));
        if (name && name_len > 0) {
            std::snprintf(name, (size_t)name_len, "%s %s", p.name, p.gcnArchName);
        }
        if (n_cu) *n_cu = p.multiProcessorCount;
        if (hbm_gb) *hbm_gb = (double)p.totalGlobalMem / 1e9;
        return 0;
    });
}

PGF_API int pgf_create(const pgf_mesh_desc* d, pgf_ctx** out) {
    pgf_ctx* c = nullptr;
    try {
        if (!d || !out) throw pgf::Error(PGF_EARG, "null descriptor");
        if (d->dim != 2 && d->dim != 3) throw pgf::Error(PGF_EARG, "dim must be 2 or 3");
        if ((d->dim == 2) != (d->elem_kind == PGF_QUAD4)) throw pgf::Error(PGF_EARG, "element kind does not match dim");
        if (d->n_mat <= 0 || d->n_mat > 255) throw pgf::Error(PGF_EARG, "n_mat must be in 1..255");
        c = new pgf_ctx();
        c->dim = d->dim;
        c->kind = d->elem_kind;
        c->n_nodes = d->n_nodes;
        c->n_all = d->n_all_nodes;
        c->n_elem = d->n_elem;
        c->n_cut = d->n_cut;
        c->n_free = d->n_free;
        c->n_pres = d->n_pres;
        c->n_rows = d->n_free + d->n_pres;
        c->n_mat = d->n_mat;
        // ---- host pattern, offsets and colours ----
        pgf::MeshView mv;
        mv.dim = d->dim; mv.n_nodes = d->n_nodes; mv.n_all = d->n_all_nodes; mv.n_elem = d->n_elem; mv.n_cut = d->n_cut;
        mv.conn = d->conn; mv.cut_elem = d->cut_elem; mv.cut_n_enrich = d->cut_n_enrich; mv.cut_enrich = d->cut_enrich;
        for (int64_t e = 0; e < d->n_elem * 4; ++e)
            if (d->conn[e] < 0 || d->conn[e] >= d->n_nodes) throw pgf::Error(PGF_EARG, "connectivity out of range");
        for (int64_t k = 0; k < d->n_cut * 16; ++k)
            if (d->cut_enrich[k] < 0 || d->n_nodes + d->cut_enrich[k] >= d->n_all_nodes)
                if (k % 16 < d->cut_n_enrich[k / 16]) throw pgf::Error(PGF_EARG, "enrichment node out of range");
        mv.finalize();
        pgf::Pattern pat;
        pat.build(mv, d->dof, d->n_free, d->n_pres);
        if (pat.nnz >= INT32_MAX) throw pgf::Error(PGF_ELIMIT, "nnz exceeds int32 range");
        c->nnz = pat.nnz;
        c->n_colors = pat.n_colors;
        c->h_rowptr = pat.rowptr;
        c->h_rowsplit = pat.rowsplit;
        c->h_colidx = pat.colidx;
        c->h_color = pat.color;
        c->nnz_ff = 0;
        for (int64_t r = 0; r < d->n_free; ++r) c->nnz_ff += pat.rowsplit[r];
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        // ---- elements in colour order ----
        std::vector<int64_t> ucount(pat.n_colors + 1, 0), ccount(pat.n_colors + 1, 0);
        for (int64_t e = 0; e < d->n_elem; ++e) {
            if (mv.cut_index[e] < 0) ++ucount[pat.color[e] + 1];
            else ++ccount[pat.color[e] + 1];
        }
        for (int k = 0; k < pat.n_colors; ++k) { ucount[k + 1] += ucount[k]; ccount[k + 1] += ccount[k]; }
        c->ue_color_ptr = ucount;
        c->ce_color_ptr = ccount;
        c->n_uncut = ucount[pat.n_colors];
        std::vector<int32_t> h_conn(4 * std::max<int64_t>(c->n_uncut, 1));
        std::vector<uint8_t> h_mat(std::max<int64_t>(c->n_uncut, 1));
        std::vector<uint32_t> h_off(8 * std::max<int64_t>(c->n_uncut, 1));
        std::vector<int32_t> h_clist(std::max<int64_t>(c->n_cut, 1));
        {
            std::vector<int64_t> uf(ucount.begin(), ucount.end() - 1), cf(ccount.begin(), ccount.end() - 1);
            for (int64_t e = 0; e < d->n_elem; ++e) {
                int col = pat.color[e];
                if (mv.cut_index[e] >= 0) { h_clist[cf[col]++] = mv.cut_index[e]; continue; }
                int64_t p = uf[col]++;
                for (int k = 0; k < 4; ++k) h_conn[k * c->n_uncut + p] = d->conn[e * 4 + k];
                if (d->elem_mat[e] < 0 || d->elem_mat[e] >= d->n_mat) throw pgf::Error(PGF_EARG, "element material out of range");
                h_mat[p] = (uint8_t)d->elem_mat[e];
                for (int k = 0; k < 8; ++k) {
                    uint32_t wv;
                    std::memcpy(&wv, &pat.elem_off[(size_t)e * 32 + k * 4], 4);
                    h_off[k * c->n_uncut + p] = wv;
                }
            }
        }
        // ---- materials ----
        std::vector<double> D(36 * d->n_mat, 0.0);
        for (int m = 0; m < d->n_mat; ++m) pgf::isotropic_D(d->dim, d->plane_strain != 0, d->mat_E[m], d->mat_nu[m], &D[36 * m]);
        std::vector<pgfk::CohParams> laws(std::max(1, d->n_coh_law));
        for (int k = 0; k < d->n_coh_law; ++k) {
            const double* pp = d->coh_params + 6 * k;
            laws[k] = pgfk::CohParams{d->coh_type[k], pp[0], pp[1], pp[2], pp[3], pp[4], pp[5]};
        }
        // ---- cut elements ----
        std::vector<int32_t> ce_nodes(16 * std::max<int64_t>(d->n_cut, 1), 0), ce_nt(std::max<int64_t>(d->n_cut, 1), 4);
        for (int64_t k = 0; k < d->n_cut; ++k) {
            int64_t e = d->cut_elem[k];
            ce_nt[k] = 4 + d->cut_n_enrich[k];
            for (int s = 0; s < ce_nt[k]; ++s) ce_nodes[k * 16 + s] = mv.node(e, s);
        }
        const int64_t n_vpt = d->n_cut ? d->vpt_ptr[d->n_cut] : 0;
        const int64_t n_cpt = d->n_cut ? d->cpt_ptr[d->n_cut] : 0;
        int64_t n_grad = 0;
        // the cut kernel stages an element's gradient blocks as one run when
        // they are consecutive (as the set-up writes them)
        c->grad_contig = true;
        for (int64_t k = 0; k < d->n_cut; ++k)
            for (int64_t pt = d->vpt_ptr[k]; pt < d->vpt_ptr[k + 1]; ++pt) {
                const int64_t m = (int64_t)ce_nt[k] * d->dim;
                if (d->vpt_grad_off[pt] != d->vpt_grad_off[d->vpt_ptr[k]] + (pt - d->vpt_ptr[k]) * m) c->grad_contig = false;
                n_grad = std::max<int64_t>(n_grad, d->vpt_grad_off[pt] + (int64_t)ce_nt[k] * d->dim);
                if (d->vpt_mat[pt] < 0 || d->vpt_mat[pt] >= d->n_mat) throw pgf::Error(PGF_EARG, "point material out of range");
            }
        for (int64_t q = 0; q < n_cpt; ++q)
            if (d->cpt_law[q] < 0 || d->cpt_law[q] >= d->n_coh_law) throw pgf::Error(PGF_EARG, "cohesive law out of range");
        for (int64_t k = 0; k < d->n_cut; ++k) {
            if (d->vpt_ptr[k + 1] - d->vpt_ptr[k] > pgfk::max_vpt(d->dim))
                throw pgf::Error(PGF_ELIMIT, "cut element with more volumetric integration points than the kernel stages");
            if (d->cpt_ptr[k + 1] - d->cpt_ptr[k] > pgfk::MAX_CPT)
                throw pgf::Error(PGF_ELIMIT, "cut element with more cohesive integration points than the kernel stages");
        }
        for (int64_t k = 0; k < d->n_cut; ++k)
            for (int64_t q = d->cpt_ptr[k]; q < d->cpt_ptr[k + 1]; ++q) {
                const int nh = d->cpt_nh[q];
                if (nh < 2 || nh > 4) throw pgf::Error(PGF_EARG, "cohesive point with a bad surface node count");
                for (int s = 0; s < 2 * nh; ++s)
                    if (d->cpt_nodes[q * 8 + s] >= ce_nt[k]) throw pgf::Error(PGF_EARG, "cohesive node slot out of range");
            }
        c->n_cpt = n_cpt;
        c->n_vpt = n_vpt;
        c->n_vgrad = n_grad;
        c->h_cut_elem.assign(d->cut_elem, d->cut_elem + d->n_cut);
        // ---- free dof -> nodal slot of the first node holding it ----
        std::vector<int64_t> fslot(std::max<int64_t>(d->n_free, 1), -1);
        for (int dd = 0; dd < d->dim; ++dd)
            for (int64_t n = 0; n < d->n_all_nodes; ++n) {
                int32_t g = d->dof[(size_t)dd * d->n_all_nodes + n];
                if (g < d->n_free && fslot[g] < 0) fslot[g] = (int64_t)dd * d->n_all_nodes + n;
            }
        // ---- upload ----
        std::vector<double> Xh(3 * d->n_all_nodes);
        std::memcpy(Xh.data(), d->coords, sizeof(double) * 3 * d->n_all_nodes);
        c->X = pgfd::dupload(Xh.data(), 3 * d->n_all_nodes);
        c->dof = pgfd::dupload(d->dof, (int64_t)d->dim * d->n_all_nodes);
        c->U = pgfd::dalloc<double>((int64_t)d->dim * d->n_all_nodes);
        HIPCHK(hipMemset(c->U, 0, sizeof(double) * d->dim * d->n_all_nodes));
        c->ue_conn = pgfd::dupload(h_conn.data(), (int64_t)h_conn.size());
        c->ue_mat = pgfd::dupload(h_mat.data(), (int64_t)h_mat.size());
        c->ue_off = pgfd::dupload(h_off.data(), (int64_t)h_off.size());
        c->Dmat = pgfd::dupload(D.data(), (int64_t)D.size());
        c->laws = pgfd::dupload(laws.data(), (int64_t)laws.size());
        c->ce_list = pgfd::dupload(h_clist.data(), (int64_t)h_clist.size());
        c->ce_nodes = pgfd::dupload(ce_nodes.data(), (int64_t)ce_nodes.size());
        c->ce_nt = pgfd::dupload(ce_nt.data(), (int64_t)ce_nt.size());
        c->vpt_ptr = pgfd::dupload(d->n_cut ? d->vpt_ptr : nullptr, d->n_cut + 1);
        c->vpt_w = pgfd::dupload(d->vpt_w, n_vpt);
        c->vpt_mat = pgfd::dupload(d->vpt_mat, n_vpt);
        c->vpt_grad_off = pgfd::dupload(d->vpt_grad_off, n_vpt);
        c->vpt_grad = pgfd::dupload(d->vpt_grad, n_grad);
        c->cpt_ptr = pgfd::dupload(d->n_cut ? d->cpt_ptr : nullptr, d->n_cut + 1);
        c->cpt_w = pgfd::dupload(d->cpt_w, n_cpt);
        c->cpt_N = pgfd::dupload(d->cpt_N, 4 * n_cpt);
        c->cpt_R = pgfd::dupload(d->cpt_R, 9 * n_cpt);
        c->cpt_nodes = pgfd::dupload(d->cpt_nodes, 8 * n_cpt);
        c->cpt_nh = pgfd::dupload(d->cpt_nh, n_cpt);
        c->cpt_law = pgfd::dupload(d->cpt_law, n_cpt);
        c->ce_off = pgfd::dupload(pat.cut_off.data(), (int64_t)pat.cut_off.size());
        c->ce_off_ptr = pgfd::dupload(pat.cut_off_ptr.data(), (int64_t)pat.cut_off_ptr.size());
        // cut elements: row-relative position of every column (bn, db) in the
        // rows of slot la, one byte each, [nt][RS] per element with the row
        // stride RS = nt * dim rounded up to 4 (rows start 4-byte aligned, so
        // the gather reads them as 32-bit words)
        {
            std::vector<int64_t> rp(d->n_cut + 1, 0);
            for (int64_t k = 0; k < d->n_cut; ++k)
                rp[k + 1] = rp[k] + (int64_t)ce_nt[k] * pgfk::cut_rel_stride(ce_nt[k] * d->dim);
            std::vector<uint8_t> rel(std::max<int64_t>(rp[d->n_cut], 1), 0);
            for (int64_t k = 0; k < d->n_cut; ++k) {
                const int nt = ce_nt[k];
                const uint8_t* off = pat.cut_off.data() + pat.cut_off_ptr[k];
                for (int la = 0; la < nt; ++la)
                    for (int bn = 0; bn < nt; ++bn) {
                        const int of = off[(la * nt + bn) * 2], op = off[(la * nt + bn) * 2 + 1];
                        const int64_t nd = ce_nodes[k * 16 + bn];
                        int cf = 0, cp = 0;
                        for (int db = 0; db < d->dim; ++db) {
                            const int64_t g = d->dof[(size_t)db * d->n_all_nodes + nd];
                            const int v = g < d->n_free ? of + cf++ : op + cp++;
                            if (v > 255) throw pgf::Error(PGF_ELIMIT, "a matrix row is wider than 256 columns");
                            rel[rp[k] + (int64_t)la * pgfk::cut_rel_stride(nt * d->dim) + bn * d->dim + db] = (uint8_t)v;
                        }
                    }
            }
            c->ce_rel = pgfd::dupload(rel.data(), (int64_t)rel.size());
            c->ce_rel_ptr = pgfd::dupload(rp.data(), d->n_cut + 1);
        }
        // gather scheme: element-order connectivity / material, incidence plan,
        // dense scratch for the cut elements' K_e and P_e
        {
            std::vector<int32_t> ec(4 * std::max<int64_t>(d->n_elem, 1));
            std::vector<uint8_t> em(std::max<int64_t>(d->n_elem, 1));
            for (int64_t e = 0; e < d->n_elem; ++e) {
                for (int k = 0; k < 4; ++k) ec[k * d->n_elem + e] = d->conn[e * 4 + k];
                em[e] = (uint8_t)d->elem_mat[e];
            }
            c->el_conn = pgfd::dupload(ec.data(), (int64_t)ec.size());
            c->el_mat = pgfd::dupload(em.data(), (int64_t)em.size());
            c->cut_index = pgfd::dupload(mv.cut_index.data(), d->n_elem);
            // each node's incidences: its uncut elements first, then its cut
            // elements, each in element order (the order in which a row's
            // entries are summed).  The gather's rounds then meet a row's cut
            // elements (whose K_e rows come from HBM) together at the end of
            // the row, not scattered between the uncut elements, so fewer
            // rounds mix the two kinds of work in one wavefront.
            {
                const size_t ni = pat.inc_elem.size();
                std::vector<int32_t> ie(ni);
                std::vector<uint8_t> is(ni);
                std::vector<uint64_t> io(ni);
                for (size_t n = 0; n + 1 < pat.inc_ptr.size(); ++n) {
                    int64_t w = pat.inc_ptr[n];
                    for (int pass = 0; pass < 2; ++pass)
                        for (int64_t k = pat.inc_ptr[n]; k < pat.inc_ptr[n + 1]; ++k)
                            if ((mv.cut_index[pat.inc_elem[k]] >= 0) == (pass == 1)) {
                                ie[w] = pat.inc_elem[k];
                                is[w] = pat.inc_slot[k];
                                io[w] = pat.inc_off[k];
                                ++w;
                            }
                }
                pat.inc_elem.swap(ie);
                pat.inc_slot.swap(is);
                pat.inc_off.swap(io);
            }
            c->n_inc = (int64_t)pat.inc_elem.size();
            c->inc_ptr = pgfd::dupload(pat.inc_ptr.data(), (int64_t)pat.inc_ptr.size());
            c->inc_elem = pgfd::dupload(pat.inc_elem.data(), c->n_inc);
            c->inc_slot = pgfd::dupload(pat.inc_slot.data(), c->n_inc);
            {
                // packed records of the gather pass: one 8-byte load per
                // incidence, one 16-byte load per element, one or two per node.
                // Uncut incidence: (element, slot | material << 16).  Cut
                // incidence: (the slot's block in the scratch, slot | 0x100 |
                // M << 9 | material << 16): the block holds the slot's K_e
                // rows, their column bytes and P_e entries (cut_slot_doubles),
                // so the gather reads them without a load of the element's
                // offsets.  The offsets follow ke_ptr, set up below.
                std::vector<int64_t> kb(d->n_cut + 1, 0);
                for (int64_t k = 0; k < d->n_cut; ++k)
                    kb[k + 1] = kb[k] + (int64_t)ce_nt[k] * pgfk::cut_slot_doubles(ce_nt[k] * d->dim, d->dim);
                if (kb[d->n_cut] > INT32_MAX) throw pgf::Error(PGF_ELIMIT, "cut-element scratch beyond 2^31 entries");
                std::vector<int32_t> rec((size_t)std::max<int64_t>(c->n_inc, 1) * 2, 0);
                for (int64_t k = 0; k < c->n_inc; ++k) {
                    const int64_t e = pat.inc_elem[k];
                    const int32_t ci = mv.cut_index[e];
                    const int32_t sl = pat.inc_slot[k], mat = (int32_t)d->elem_mat[e] << 16;
                    if (ci >= 0) {
                        const int m = ce_nt[ci] * d->dim;
                        rec[2 * k] = (int32_t)(kb[ci] + (int64_t)sl * pgfk::cut_slot_doubles(m, d->dim));
                        rec[2 * k + 1] = sl | 0x100 | (m << 9) | mat;
                    } else {
                        rec[2 * k] = (int32_t)e;
                        rec[2 * k + 1] = sl | mat;
                    }
                }
                c->inc_rec = pgfd::dupload(rec.data(), (int64_t)rec.size());
                c->el_conn4 = pgfd::dupload(d->conn, 4 * std::max<int64_t>(d->n_elem, 1));
                const int QW = d->dim == 2 ? 2 : 4;
                std::vector<double> xq((size_t)QW * d->n_all_nodes, 0.0);
                for (int64_t n = 0; n < d->n_all_nodes; ++n)
                    for (int i = 0; i < d->dim; ++i) xq[(size_t)n * QW + i] = d->coords[(size_t)i * d->n_all_nodes + n];
                c->Xq = pgfd::dupload(xq.data(), (int64_t)xq.size());
            }
            // per incidence of an uncut element: the row-relative position of
            // each of the element's 4 * dim columns (free part, then
            // prescribed part of the row), one byte each, so the gather pass
            // needs no dof lookups to place its entries
            {
                const int W = 4 * d->dim;
                std::vector<uint8_t> rel((size_t)std::max<int64_t>(c->n_inc, 1) * W, 0);
                bool wide = false;
                for (int64_t k = 0; k < c->n_inc; ++k) {
                    const int64_t e = pat.inc_elem[k];
                    if (mv.cut_index[e] >= 0) continue;
                    const uint64_t off = pat.inc_off[k];
                    for (int bn = 0; bn < 4; ++bn) {
                        const int of = (int)((off >> (16 * bn)) & 0xff), op = (int)((off >> (16 * bn + 8)) & 0xff);
                        const int64_t nd = d->conn[e * 4 + bn];
                        int cf = 0, cp = 0;
                        for (int db = 0; db < d->dim; ++db) {
                            const int64_t g = d->dof[(size_t)db * d->n_all_nodes + nd];
                            const int v = g < d->n_free ? of + cf++ : op + cp++;
                            if (v > 255) wide = true;
                            rel[(size_t)k * W + bn * d->dim + db] = (uint8_t)v;
                        }
                    }
                }
                if (wide) throw pgf::Error(PGF_ELIMIT, "a matrix row is wider than 256 columns");
                c->inc_rel = pgfd::dupload(rel.data(), (int64_t)rel.size());
            }
            std::vector<int64_t> kp(d->n_cut + 1, 0), pp(d->n_cut + 1, 0);
            for (int64_t k = 0; k < d->n_cut; ++k) {
                const int64_t m = (int64_t)ce_nt[k] * d->dim;
                kp[k + 1] = kp[k] + (int64_t)ce_nt[k] * pgfk::cut_slot_doubles((int)m, d->dim);
                pp[k + 1] = pp[k] + m;
            }
            c->n_ke = kp[d->n_cut];
            c->ke_ptr = pgfd::dupload(kp.data(), d->n_cut + 1);
            c->pe_ptr = pgfd::dupload(pp.data(), d->n_cut + 1);
            c->ke = pgfd::dalloc<double>(kp[d->n_cut]);
            if (d->n_cut > 0) {
                hipLaunchKernelGGL(pgfd::k_place_rel, dim3((unsigned)((d->n_cut + 255) / 256)), dim3(256), 0, c->stream,
                                   d->n_cut, d->dim, c->ce_nt, c->ce_rel, c->ce_rel_ptr, c->ke_ptr, c->ke);
                HIPCHK(hipGetLastError());
                HIPCHK(hipStreamSynchronize(c->stream));
            }
            c->pe = pgfd::dalloc<double>(pp[d->n_cut]);
            c->h_dof.assign(d->dof, d->dof + (size_t)d->dim * d->n_all_nodes);
            if (pat.shared_dofs) {
                // periodic images: nodes whose dofs an earlier node holds
                std::vector<int32_t> first((size_t)c->n_rows, -1);
                c->h_image.assign(d->n_all_nodes, 0);
                for (int64_t n = 0; n < d->n_all_nodes; ++n) {
                    const int32_t g = d->dof[n];
                    if (first[g] < 0) first[g] = (int32_t)n;
                    else c->h_image[n] = 1;
                }
                c->image = pgfd::dupload(c->h_image.data(), d->n_all_nodes);
            }
            c->h_inc_ptr = pat.inc_ptr;
            pgfd::build_gather_blocks(*c, pgfd::default_gather_slots(d->dim), false);
        }
        c->dmax_conv = pgfd::dalloc<double>(n_cpt);
        c->dmax_trial = pgfd::dalloc<double>(n_cpt);
        c->flags = pgfd::dalloc<uint8_t>(n_cpt);
        HIPCHK(hipMemset(c->dmax_conv, 0, sizeof(double) * std::max<int64_t>(n_cpt, 1)));
        HIPCHK(hipMemset(c->dmax_trial, 0, sizeof(double) * std::max<int64_t>(n_cpt, 1)));
        HIPCHK(hipMemset(c->flags, 0, std::max<int64_t>(n_cpt, 1)));
        c->free_slot = pgfd::dupload(fslot.data(), (int64_t)fslot.size());
        std::vector<int32_t> rp32(c->n_rows + 1);
        for (int64_t r = 0; r <= c->n_rows; ++r) rp32[r] = (int32_t)pat.rowptr[r];
        c->rowptr = pgfd::dupload(rp32.data(), c->n_rows + 1);
        c->rowsplit = pgfd::dupload(pat.rowsplit.data(), c->n_rows);
        // node groups of the free rows for the SpMV: runs of at most three
        // consecutive rows with the same free column list
        {
            std::vector<int32_t> grp{0};
            for (int64_t r = 1; r < c->n_free; ++r) {
                const int64_t r0 = grp.back();
                bool same = r - r0 < 3 && pat.rowsplit[r] == pat.rowsplit[r0];
                if (same)
                    same = std::equal(pat.colidx.begin() + pat.rowptr[r], pat.colidx.begin() + pat.rowptr[r] + pat.rowsplit[r],
                                      pat.colidx.begin() + pat.rowptr[r0]);
                if (!same) grp.push_back((int32_t)r);
            }
            if (c->n_free > 0) grp.push_back((int32_t)c->n_free);
            c->n_grp = c->n_free > 0 ? (int64_t)grp.size() - 1 : 0;
            c->grp_row = pgfd::dupload(grp.data(), (int64_t)grp.size());
        }
        c->colidx = pgfd::dupload(pat.colidx.data(), c->nnz);
        c->K = pgfd::dalloc<double>(c->nnz);
        c->K32 = pgfd::dalloc<float>(c->nnz);
        c->P = pgfd::dalloc<double>(c->n_rows);
        c->Fext_full = pgfd::dalloc<double>(c->n_free);
        c->upres_full = pgfd::dalloc<double>(c->n_pres);
        HIPCHK(hipMemset(c->Fext_full, 0, sizeof(double) * std::max<int64_t>(c->n_free, 1)));
        HIPCHK(hipMemset(c->upres_full, 0, sizeof(double) * std::max<int64_t>(c->n_pres, 1)));
        const int64_t nf = c->n_free;
        c->r = pgfd::dalloc<double>(nf); c->x = pgfd::dalloc<double>(nf); c->z = pgfd::dalloc<double>(nf);
        c->p0 = pgfd::dalloc<double>(nf); c->p1 = pgfd::dalloc<double>(nf); c->q = pgfd::dalloc<double>(nf);
        c->dinv = pgfd::dalloc<double>(nf); c->guess = pgfd::dalloc<double>(nf); c->res = pgfd::dalloc<double>(nf);
        c->partial = pgfd::dalloc<double>(std::max(3 * 2 * pgfd::RED_BLOCKS, pgfd::CG_PART));
        c->scal = pgfd::dalloc<double>(pgfd::S_NSCAL + 8);
        c->counter = pgfd::dalloc<unsigned>(8);
        c->done = pgfd::dalloc<int>(1);
        HIPCHK(hipMemset(c->counter, 0, sizeof(unsigned) * 8));
        HIPCHK(hipMemset(c->done, 0, sizeof(int)));
        HIPCHK(hipMemset(c->scal, 0, sizeof(double) * (pgfd::S_NSCAL + 8)));
        HIPCHK(hipHostMalloc((void**)&c->h_scal, sizeof(double) * (pgfd::S_NSCAL + 8), hipHostMallocDefault));
        HIPCHK(hipDeviceSynchronize());
        *out = c;
        return 0;
    } catch (const pgf::Error& e) {
        pgfd::g_last_error = e.what();
        delete c;
        return e.code;
    } catch (const std::exception& e) {
        pgfd::g_last_error = e.what();
        delete c;
        return PGF_EARG;
    }
}

PGF_API void pgf_destroy(pgf_ctx* ctx) { delete ctx; }

PGF_API int pgf_csr_info_get(pgf_ctx* c, pgf_csr_info* info) {
    if (!c || !info) return PGF_EARG;
    info->n_rows = c->n_rows; info->n_free = c->n_free; info->nnz = c->nnz; info->nnz_ff = c->nnz_ff;
    info->n_colors = c->n_colors; info->reserved = 0;
    return 0;
}

PGF_API int pgf_get_csr(pgf_ctx* c, int64_t* rowptr, int32_t* rowsplit, int32_t* colidx) {
    if (!c) return PGF_EARG;
    if (rowptr) std::memcpy(rowptr, c->h_rowptr.data(), sizeof(int64_t) * (c->n_rows + 1));
    if (rowsplit) std::memcpy(rowsplit, c->h_rowsplit.data(), sizeof(int32_t) * c->n_rows);
    if (colidx) std::memcpy(colidx, c->h_colidx.data(), sizeof(int32_t) * c->nnz);
    return 0;
}

PGF_API int pgf_get_colors(pgf_ctx* c, int32_t* color) {
    if (!c || !color) return PGF_EARG;
    std::memcpy(color, c->h_color.data(), sizeof(int32_t) * c->n_elem);
    return 0;
}

PGF_API int pgf_set_solution(pgf_ctx* c, const double* u) {
    if (!c || !u) return PGF_EARG;
    return guard(c, [&]() -> int {
        HIPCHK(hipMemcpyAsync(c->U, u, sizeof(double) * c->dim * c->n_all, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return 0;
    });
}

PGF_API int pgf_set_solution_dofs(pgf_ctx* c, const double* x_free, const double* u_pres) {
    if (!c || !x_free) return PGF_EARG;
    return guard(c, [&]() -> int {
        double* dx = pgfd::dupload(x_free, std::max<int64_t>(c->n_free, 1));
        double* dp = u_pres ? pgfd::dupload(u_pres, std::max<int64_t>(c->n_pres, 1)) : nullptr;
        pgfd::enqueue_set_from_dofs(*c, dx, dp);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(dx));
        if (dp) HIPCHK(hipFree(dp));
        return 0;
    });
}

PGF_API int pgf_volume_average(pgf_ctx* c, const double* eps_bar, int32_t n_seg, const int64_t* seg_elem, double* out) {
    if (!c || !eps_bar || n_seg <= 0 || !seg_elem || !out) return PGF_EARG;
    return guard(c, [&]() -> int {
        for (int32_t k = 0; k < n_seg; ++k)
            if (seg_elem[k] < 0 || seg_elem[k] > seg_elem[k + 1] || seg_elem[k + 1] > c->n_elem)
                throw pgf::Error(PGF_EARG, "element segments must be ordered ranges inside [0, n_elem]");
        const int S = c->dim == 2 ? 3 : 6, W = 2 * S + 1;
        double* de = pgfd::dupload(eps_bar, S);
        int64_t* ds = pgfd::dupload(seg_elem, (int64_t)n_seg + 1);
        double* dpart = pgfd::dalloc<double>((int64_t)n_seg * 64 * W);
        double* dout = pgfd::dalloc<double>((int64_t)n_seg * W);
        pgfd::enqueue_volume_average(*c, de, n_seg, ds, dpart, dout);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * n_seg * W, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (void* p : {(void*)de, (void*)ds, (void*)dpart, (void*)dout}) HIPCHK(hipFree(p));
        return 0;
    });
}

PGF_API int pgf_volume_average_noalloc(pgf_ctx* c, const double* eps_bar, int32_t n_seg, const int64_t* seg_elem,
                                      double* out) {
    if (!c || !eps_bar || n_seg <= 0 || !seg_elem || !out) return PGF_EARG;
    if (n_seg > PGF_NOALLOC_MAX_SEG) {
        c->err = "pgf_volume_average_noalloc: at most " + std::to_string(PGF_NOALLOC_MAX_SEG) + " segments";
        return PGF_ELIMIT;
    }
    return guard(c, [&]() -> int {
        for (int32_t k = 0; k < n_seg; ++k)
            if (seg_elem[k] < 0 || seg_elem[k] > seg_elem[k + 1] || seg_elem[k + 1] > c->n_elem)
                throw pgf::Error(PGF_EARG, "element segments must be ordered ranges inside [0, n_elem]");
        double* base = pgfd::noalloc_scratch();
        const int S = c->dim == 2 ? 3 : 6, W = 2 * S + 1;
        // scratch layout: eps_bar [8] | segments [MAX_SEG + 1] int64 | out [MAX_SEG][13] | partials
        double* de = base;
        int64_t* ds = (int64_t*)(base + 8);
        double* dout = base + 8 + (PGF_NOALLOC_MAX_SEG + 1);
        double* dpart = dout + (size_t)PGF_NOALLOC_MAX_SEG * 13;
        HIPCHK(hipMemcpyAsync(de, eps_bar, sizeof(double) * S, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(ds, seg_elem, sizeof(int64_t) * (n_seg + 1), hipMemcpyHostToDevice, c->stream));
        pgfd::enqueue_volume_average(*c, de, n_seg, ds, dpart, dout);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * n_seg * W, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return 0;
    });
}

PGF_API int pgf_get_solution(pgf_ctx* c, double* u) {
    if (!c || !u) return PGF_EARG;
    return guard(c, [&]() -> int {
        HIPCHK(hipMemcpyAsync(u, c->U, sizeof(double) * c->dim * c->n_all, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return 0;
    });
}

PGF_API int pgf_set_load(pgf_ctx* c, const double* up, const double* fext) {
    if (!c) return PGF_EARG;
    return guard(c, [&]() -> int {
        if (up && c->n_pres)
            HIPCHK(hipMemcpyAsync(c->upres_full, up, sizeof(double) * c->n_pres, hipMemcpyHostToDevice, c->stream));
        if (fext && c->n_free)
            HIPCHK(hipMemcpyAsync(c->Fext_full, fext, sizeof(double) * c->n_free, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return 0;
    });
}

PGF_API int pgf_apply_load_fraction(pgf_ctx* c, double frac) {
    if (!c) return PGF_EARG;
    return guard(c, [&]() -> int {
        if (c->n_pres)
            hipLaunchKernelGGL(pgfd::k_apply_frac, dim3(pgfd::grid_for(c->dim * c->n_all)), dim3(256), 0, c->stream,
                               (int64_t)c->dim * c->n_all, c->dof, c->n_free, c->upres_full, frac, c->U);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        return 0;
    });
}

PGF_API int pgf_assemble(pgf_ctx* c, double dt, uint32_t flags) {
    if (!c) return PGF_EARG;
    return guard(c, [&]() -> int {
        pgfd::enqueue_assembly(*c, dt, flags);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
        return 0;
    });
}

PGF_API int pgf_get_matrix(pgf_ctx* c, double* vals) {
    if (!c || !vals) return PGF_EARG;
    return guard(c, [&]() -> int {
        HIPCHK(hipMemcpy(vals, c->K, sizeof(double) * c->nnz, hipMemcpyDeviceToHost));
        return 0;
    });
}
PGF_API int pgf_get_matrix32(pgf_ctx* c, float* vals) {
    if (!c || !vals) return PGF_EARG;
    return guard(c, [&]() -> int {
        HIPCHK(hipMemcpy(vals, c->K32, sizeof(float) * c->nnz, hipMemcpyDeviceToHost));
        return 0;
    });
}
PGF_API int pgf_get_internal_force(pgf_ctx* c, double* p) {
    if (!c || !p) return PGF_EARG;
    return guard(c, [&]() -> int {
        HIPCHK(hipMemcpy(p, c->P, sizeof(double) * c->n_rows, hipMemcpyDeviceToHost));
        return 0;
    });
}

PGF_API int pgf_n_cohesive_points(pgf_ctx* c, int64_t* n) {
    if (!c || !n) return PGF_EARG;
    *n = c->n_cpt;
    return 0;
}
PGF_API int pgf_isv_commit(pgf_ctx* c) {
    if (!c) return PGF_EARG;
    return guard(c, [&]() -> int {
        if (c->n_cpt)
            hipLaunchKernelGGL(pgfd::k_copy, dim3(pgfd::grid_for(c->n_cpt)), dim3(256), 0, c->stream, c->n_cpt, c->dmax_trial,
                               c->dmax_conv);
        HIPCHK(hipStreamSynchronize(c->stream));
        return 0;
    });
}
PGF_API int pgf_isv_rollback(pgf_ctx* c) {
    if (!c) return PGF_EARG;
    return guard(c, [&]() -> int {
        if (c->n_cpt)
            hipLaunchKernelGGL(pgfd::k_copy, dim3(pgfd::grid_for(c->n_cpt)), dim3(256), 0, c->stream, c->n_cpt, c->dmax_conv,
                               c->dmax_trial);
        HIPCHK(hipStreamSynchronize(c->stream));
        return 0;
    });
}
PGF_API int pgf_get_isv(pgf_ctx* c, double* dt, double* dc, uint8_t* fl) {
    if (!c) return PGF_EARG;
    return guard(c, [&]() -> int {
        if (c->n_cpt) {
            if (dt) HIPCHK(hipMemcpy(dt, c->dmax_trial, sizeof(double) * c->n_cpt, hipMemcpyDeviceToHost));
            if (dc) HIPCHK(hipMemcpy(dc, c->dmax_conv, sizeof(double) * c->n_cpt, hipMemcpyDeviceToHost));
            if (fl) HIPCHK(hipMemcpy(fl, c->flags, c->n_cpt, hipMemcpyDeviceToHost));
        }
        return 0;
    });
}
PGF_API int pgf_set_isv(pgf_ctx* c, const double* dc) {
    if (!c || !dc) return PGF_EARG;
    return guard(c, [&]() -> int {
        if (c->n_cpt) {
            HIPCHK(hipMemcpy(c->dmax_conv, dc, sizeof(double) * c->n_cpt, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(c->dmax_trial, dc, sizeof(double) * c->n_cpt, hipMemcpyHostToDevice));
        }
        return 0;
    });
}

PGF_API int pgf_spmv(pgf_ctx* c, const double* x, double* y) {
    if (!c || !x || !y) return PGF_EARG;
    return guard(c, [&]() -> int {
        HIPCHK(hipMemcpyAsync(c->x, x, sizeof(double) * c->n_free, hipMemcpyHostToDevice, c->stream));
        pgfd::enqueue_spmv<double, double>(*c, c->K, c->x, c->q);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(y, c->q, sizeof(double) * c->n_free, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return 0;
    });
}

PGF_API int pgf_pcg(pgf_ctx* c, const double* b, double* x, const double* x0, double rtol, int32_t maxit,
                    int32_t precision, pgf_solve_info* info) {
    if (!c || !b || !x) return PGF_EARG;
    return guard(c, [&]() -> int {
        const int64_t n = c->n_free;
        HIPCHK(hipMemcpyAsync(c->res, b, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        if (x0) HIPCHK(hipMemcpyAsync(c->x, x0, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        // precision 32: FP32 matrix values (half the value bytes per SpMV),
        // FP64 Krylov vectors and recurrences
        if (precision == 32)
            pgfd::run_pcg<float, double>(*c, c->K32, c->res, true, c->x, c->r, c->z, c->p0, c->p1, c->q, c->dinv,
                                         x0 != nullptr, rtol, maxit, info);
        else
            pgfd::run_pcg<double, double>(*c, c->K, c->res, true, c->x, c->r, c->z, c->p0, c->p1, c->q, c->dinv,
                                          x0 != nullptr, rtol, maxit, info);
        HIPCHK(hipMemcpy(x, c->x, sizeof(double) * n, hipMemcpyDeviceToHost));
        return 0;
    });
}

PGF_API int pgf_set_guess(pgf_ctx* c, const double* dx) {
    if (!c) return PGF_EARG;
    return guard(c, [&]() -> int {
        if (!dx) { c->have_guess = false; return 0; }
        HIPCHK(hipMemcpy(c->guess, dx, sizeof(double) * c->n_free, hipMemcpyHostToDevice));
        c->have_guess = true;
        return 0;
    });
}

// Newton-Raphson for one load step, SolverNonlinearKSP::nonlinearStep
// (src/SolverNonlinearKSP.cpp:55-159): K and r at the current iterate, linear
// solve, update, new residual; converged when ||r|| < rel_tol ||r0|| or
// ||r|| < abs_tol, failed on NaN, iteration cap, or three increases in a row.
PGF_API int pgf_newton_step(pgf_ctx* c, double frac, const pgf_newton_opts* o, pgf_step_info* info) {
    if (!c || !o) return PGF_EARG;
    return guard(c, [&]() -> int {
        hipEvent_t t0, t1, a0, a1;
        HIPCHK(hipEventCreate(&t0)); HIPCHK(hipEventCreate(&t1));
        HIPCHK(hipEventCreate(&a0)); HIPCHK(hipEventCreate(&a1));
        HIPCHK(hipEventRecord(t0, c->stream));
        const int64_t nf = c->n_free;
        double ms_asm = 0.0, ms_sol = 0.0;
        if (c->n_pres)
            hipLaunchKernelGGL(pgfd::k_apply_frac, dim3(pgfd::grid_for(c->dim * c->n_all)), dim3(256), 0, c->stream,
                               (int64_t)c->dim * c->n_all, c->dof, c->n_free, c->upres_full, frac, c->U);
        auto residual = [&](double& rn) {
            HIPCHK(hipEventRecord(a0, c->stream));
            pgfd::enqueue_assembly(*c, o->dt, PGF_ASM_FUNC);
            hipLaunchKernelGGL(pgfd::k_residual, dim3(pgfd::grid_for(nf)), dim3(256), 0, c->stream, nf, c->Fext_full, frac,
                               c->P, c->r, c->partial, c->counter + 3, c->scal + pgfd::S_NSCAL);
            HIPCHK(hipMemcpyAsync(c->h_scal + pgfd::S_NSCAL, c->scal + pgfd::S_NSCAL, sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
            HIPCHK(hipEventRecord(a1, c->stream));
            HIPCHK(hipEventSynchronize(a1));
            float ms;
            HIPCHK(hipEventElapsedTime(&ms, a0, a1));
            ms_asm += ms;
            rn = std::sqrt(c->h_scal[pgfd::S_NSCAL]);
        };
        double res0, res;
        residual(res0);
        res = res0;
        int it = 0, inc = 0, pcg_its = 0, fallback = 0;
        bool conv = res0 < o->abs_tol || res0 == 0.0, fail = false;
        while (!conv && !fail) {
            const bool fp32 = o->precision == 32 && fallback == 0;
            HIPCHK(hipEventRecord(a0, c->stream));
            pgfd::enqueue_assembly(*c, o->dt, PGF_ASM_JAC | (fp32 ? PGF_ASM_FP32 : 0u));
            HIPCHK(hipEventRecord(a1, c->stream));
            HIPCHK(hipEventSynchronize(a1));
            float ms;
            HIPCHK(hipEventElapsedTime(&ms, a0, a1));
            ms_asm += ms;
            pgf_solve_info si{};
            // inexact Newton: solve only as far as this iteration needs to reach the target
            double lin_tol = o->lin_rtol;
            if (o->adaptive_lin && res > 0.0)
                lin_tol = std::min(0.5, std::max(o->lin_rtol, 0.5 * o->rel_tol * res0 / res));
            bool g = o->use_guess == 1 && c->have_guess && it == 0;
            if (c->learn && (it == 0 || o->use_guess == 2 || o->use_guess == 3))
                g = pgfd::learn_guess(*c, o->use_guess, fp32, it) || g;
            if (g) hipLaunchKernelGGL(pgfd::k_copy, dim3(pgfd::grid_for(nf)), dim3(256), 0, c->stream, nf, c->guess, c->x);
            // FP32 mode: FP32 tangent values with FP64 Krylov vectors; the
            // residual, the update and delta_max stay FP64 either way.  When
            // the FP32 matrix cannot reach the tolerance (rounding K to FP32
            // perturbs it by ~6e-8 relative, which a Jacobi-scaled condition
            // number near 1e7 turns into stagnation), the solve continues
            // from its iterate on the FP64 matrix and the step stays in FP64.
            if (fp32) {
                hipLaunchKernelGGL(pgfd::k_copy, dim3(pgfd::grid_for(nf)), dim3(256), 0, c->stream, nf, c->r, c->res);
                const int maxit32 = o->fp32_maxit > 0 ? std::min(o->fp32_maxit, o->lin_maxit) : o->lin_maxit;
                pgfd::run_pcg<float, double>(*c, c->K32, c->res, true, c->x, c->r, c->z, c->p0, c->p1, c->q, c->dinv, g,
                                             lin_tol, maxit32, &si);
                if (!si.converged) {
                    fallback = it + 1;
                    HIPCHK(hipEventRecord(a0, c->stream));
                    pgfd::enqueue_assembly(*c, o->dt, PGF_ASM_JAC);
                    HIPCHK(hipEventRecord(a1, c->stream));
                    HIPCHK(hipEventSynchronize(a1));
                    HIPCHK(hipEventElapsedTime(&ms, a0, a1));
                    ms_asm += ms;
                    const bool warm = std::isfinite(si.rel_residual) && si.rel_residual < 1.0;
                    pgf_solve_info s64{};
                    pgfd::run_pcg<double, double>(*c, c->K, c->res, true, c->x, c->r, c->z, c->p0, c->p1, c->q, c->dinv,
                                                  warm, lin_tol, std::max(o->lin_maxit - si.iterations, 1), &s64);
                    si.iterations += s64.iterations;
                    si.ms += s64.ms;
                    si.converged = s64.converged;
                    si.rel_residual = s64.rel_residual;
                }
            } else
                pgfd::run_pcg<double, double>(*c, c->K, c->r, true, c->x, c->r, c->z, c->p0, c->p1, c->q, c->dinv, g,
                                              lin_tol, o->lin_maxit, &si);
            hipLaunchKernelGGL((pgfd::k_update_U<double>), dim3(pgfd::grid_for(c->dim * c->n_all)), dim3(256), 0,
                               c->stream, (int64_t)c->dim * c->n_all, c->dof, nf, c->x, c->U);
            ms_sol += si.ms;
            pcg_its += si.iterations;
            ++it;
            double prev = res;
            residual(res);
            if (res < o->rel_tol * res0 || res < o->abs_tol) { conv = true; break; }
            if (!std::isfinite(res)) { fail = true; break; }
            if (it >= o->max_iter) { fail = true; break; }
            inc = res > prev ? inc + 1 : 0;
            if (inc >= 3) { fail = true; break; }
        }
        HIPCHK(hipEventRecord(t1, c->stream));
        HIPCHK(hipEventSynchronize(t1));
        float mt;
        HIPCHK(hipEventElapsedTime(&mt, t0, t1));
        (void)hipEventDestroy(t0); (void)hipEventDestroy(t1); (void)hipEventDestroy(a0); (void)hipEventDestroy(a1);
        if (info) {
            info->newton_iterations = it;
            info->converged = conv ? 1 : 0;
            info->pcg_iterations = pcg_its;
            info->fp64_fallback = fallback;
            info->res0 = res0;
            info->res = res;
            info->ms_total = mt;
            info->ms_assemble = ms_asm;
            info->ms_solve = ms_sol;
        }
        return conv ? 0 : PGF_ENOTCONVERGED;
    });
}

PGF_API int pgf_time_assemble(pgf_ctx* c, uint32_t flags, int32_t reps, double* ms_per_call) {
    if (!c || reps <= 0 || !ms_per_call) return PGF_EARG;
    return guard(c, [&]() -> int {
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1));
        pgfd::enqueue_assembly(*c, 1.0, flags);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipEventRecord(e0, c->stream));
        for (int i = 0; i < reps; ++i) pgfd::enqueue_assembly(*c, 1.0, flags);
        HIPCHK(hipEventRecord(e1, c->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_call = ms / reps;
        (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
        return 0;
    });
}

// Per-pass device time of the gather assembly, averaged over `reps` calls:
// out[0] whole call, out[1] cut-element pass, out[2] row-gather pass.
PGF_API int pgf_time_assemble_split(pgf_ctx* c, uint32_t flags, int32_t reps, double* out) {
    if (!c || reps <= 0 || !out || (flags & 16u)) return PGF_EARG;
    return guard(c, [&]() -> int {
        hipEvent_t ev[3], e0, e1;
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1));
        double cut = 0.0, gat = 0.0;
        pgfd::enqueue_assembly(*c, 1.0, flags);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipEventRecord(e0, c->stream));
        for (int i = 0; i < reps; ++i) {
            c->timing_events = ev;
            pgfd::enqueue_assembly(*c, 1.0, flags);
            c->timing_events = nullptr;
            HIPCHK(hipEventSynchronize(ev[2]));
            float a, b;
            HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
            HIPCHK(hipEventElapsedTime(&b, ev[1], ev[2]));
            cut += a;
            gat += b;
        }
        HIPCHK(hipEventRecord(e1, c->stream));
        HIPCHK(hipEventSynchronize(e1));
        float tot;
        HIPCHK(hipEventElapsedTime(&tot, e0, e1));
        out[0] = tot / reps;
        out[1] = cut / reps;
        out[2] = gat / reps;
        for (auto& e : ev) (void)hipEventDestroy(e);
        (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
        return 0;
    });
}

PGF_API int pgf_set_gather_layout(pgf_ctx* c, int32_t slots, int32_t node_rows) {
    if (!c || (node_rows != 0 && node_rows != 1)) return PGF_EARG;
    return guard(c, [&]() -> int {
        HIPCHK(hipStreamSynchronize(c->stream));
        pgfd::build_gather_blocks(*c, slots, node_rows != 0);
        return 0;
    });
}

PGF_API int pgf_get_gather_layout(pgf_ctx* c, int32_t* slots, int32_t* node_rows) {
    if (!c || !slots || !node_rows) return PGF_EARG;
    *slots = c->gslots;
    *node_rows = c->gnode ? 1 : 0;
    return 0;
}

PGF_API int pgf_time_spmv(pgf_ctx* c, int32_t precision, int32_t reps, double* ms_per_call) {
    if (!c || reps <= 0 || !ms_per_call) return PGF_EARG;
    return guard(c, [&]() -> int {
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0)); HIPCHK(hipEventCreate(&e1));
        // precision 32 times the solver's mixed product: FP32 values, FP64 vectors
        pgfd::fill_async<double>(*c, c->x, c->n_free, 1.0);
        if (precision == 32) pgfd::enqueue_spmv<float, double>(*c, c->K32, c->x, c->q);
        else pgfd::enqueue_spmv<double, double>(*c, c->K, c->x, c->q);
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipEventRecord(e0, c->stream));
        for (int i = 0; i < reps; ++i) {
            if (precision == 32) pgfd::enqueue_spmv<float, double>(*c, c->K32, c->x, c->q);
            else pgfd::enqueue_spmv<double, double>(*c, c->K, c->x, c->q);
        }
        HIPCHK(hipEventRecord(e1, c->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_call = ms / reps;
        (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
        return 0;
    });
}

}  // extern "C"
