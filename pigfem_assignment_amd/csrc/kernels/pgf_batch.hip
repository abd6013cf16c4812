// Batched solves of many independent systems sharing one device context: the
// free-free block K_ff of a context whose mesh is a concatenation of
// independent meshes (RVEs) is block diagonal, one block per RVE, and each
// block is solved for NR right-hand sides (load cases / sensitivity
// right-hand sides) at once.
//
//  * segmented Jacobi-PCG: every (system, rhs) pair runs its own CG
//    recurrence (its own alpha, beta and convergence test), all pairs advance
//    together, one SpMM sweep per iteration reads each K entry once for all NR
//    vectors; row tiles never cross a system boundary, tiles of converged
//    systems are skipped, so the work is the sum over systems of their own
//    iteration counts; when the live tiles fall below 60 % of those launched
//    the launch list is compacted to the live ones (and the graph captured
//    again), so converged systems stop costing workgroup launches;
//  * a chunk of iterations is captured once in a hipGraph and replayed until
//    every pair has converged (or hit maxit); clearing is done by kernels, the
//    graph holds no memset nodes;
//  * the tile size (rows per tile) and threads per row are chosen by a small
//    autotuner that times a few iterations of each candidate on the batch.
// Vectors are interleaved [n_free][NR].
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "../../../include/pgfem.h"
#include "pgf_device.hpp"

namespace pgfb {

constexpr int BT = 256;  // threads per tile workgroup

struct BArgs {
    int64_t n_tiles;
    int64_t n_launch;          // tiles launched per sweep (the live ones)
    const int32_t* tile_list;  // [n_launch] tile of each launched workgroup
    const int32_t* tile_sys;   // [n_tiles]
    const int64_t* tile_row;   // [n_tiles + 1] row range of each tile
    const int32_t* rowptr;
    const int32_t* rowsplit;
    const int32_t* colidx;
    const double* K;
    const double* dinv;        // [n_free]
    double *x, *r, *z, *pnew, *q;  // [n_free][NR]; pnew is p, updated in place
    double* tpart;             // [n_tiles][NR]
    double* tpart2;            // [n_tiles][2*NR]
    const int64_t* sys_tile;   // [n_sys + 1] tile range of each system
    double* scal;              // [n_sys][NR][8]: rz, pq, alpha, beta, rr, rr0, tol2, iters
    int* done;                 // [n_sys][NR]
    int* active;               // [1] nonzero while some pair is still running
    int n_sys;
};
enum { S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_RR0, S_TOL2, S_IT, S_N };

template <int NR>
__device__ inline bool tile_done(const BArgs& a, int s) {
    bool all = true;
    for (int c = 0; c < NR; ++c) all &= a.done[s * NR + c] != 0;
    return all;
}

template <int NR>
__device__ inline void block_sum(double (&v)[NR], double* out) {
    __shared__ double sh[BT / 64][NR];
    for (int c = 0; c < NR; ++c)
        for (int m = 32; m >= 1; m >>= 1) v[c] += __shfl_xor(v[c], m, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int c = 0; c < NR; ++c) sh[w][c] = v[c];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int c = 0; c < NR; ++c) {
            double s = 0.0;
            for (int k = 0; k < BT / 64; ++k) s += sh[k][c];
            out[c] = s;
        }
    __syncthreads();
}

// p = z + beta p on the tile's rows, in place, for the pairs still running
// (li: index into the launch list; the whole block calls it)
template <int NR>
__device__ inline void bpupd_tile(const BArgs& a, int64_t li) {
    const int64_t t = a.tile_list[li];
    const int s = a.tile_sys[t];
    if (tile_done<NR>(a, s)) return;
    double beta[NR];
    bool act[NR];
    for (int c = 0; c < NR; ++c) {
        beta[c] = a.scal[(s * NR + c) * S_N + S_BETA];
        act[c] = a.done[s * NR + c] == 0;
    }
    for (int64_t r = a.tile_row[t] + threadIdx.x; r < a.tile_row[t + 1]; r += BT)
        for (int c = 0; c < NR; ++c)
            if (act[c]) a.pnew[r * NR + c] = a.z[r * NR + c] + beta[c] * a.pnew[r * NR + c];
}
template <int NR>
__global__ __launch_bounds__(BT) void k_bpupd(BArgs a) {
    bpupd_tile<NR>(a, blockIdx.x);
}

// q = K p over the row's columns (p gathered once per column for all NR
// right-hand sides), partial p . q.  G threads per row.
template <int NR, int G>
__device__ inline void bspmm_tile(const BArgs& a, int64_t li) {
    const int64_t t = a.tile_list[li];
    const int s = a.tile_sys[t];
    if (tile_done<NR>(a, s)) return;
    bool act[NR];
    for (int c = 0; c < NR; ++c) act[c] = a.done[s * NR + c] == 0;
    const int64_t r0 = a.tile_row[t], r1 = a.tile_row[t + 1];
    const int sub = threadIdx.x % G;
    double pq[NR];
    for (int c = 0; c < NR; ++c) pq[c] = 0.0;
    for (int64_t r = r0 + threadIdx.x / G; r < r1; r += BT / G) {
        double acc[NR];
        for (int c = 0; c < NR; ++c) acc[c] = 0.0;
        const int32_t b = a.rowptr[r], e = b + a.rowsplit[r];
        for (int32_t j = b + sub; j < e; j += G) {
            const double v = a.K[j];
            const int64_t col = a.colidx[j];
            for (int c = 0; c < NR; ++c) acc[c] += v * a.pnew[col * NR + c];
        }
        for (int c = 0; c < NR; ++c)
            for (int m = G / 2; m >= 1; m >>= 1) acc[c] += __shfl_xor(acc[c], m, G);
        if (sub == 0)
            for (int c = 0; c < NR; ++c) {
                if (!act[c]) continue;
                a.q[r * NR + c] = acc[c];
                pq[c] += a.pnew[r * NR + c] * acc[c];
            }
    }
    double tot[NR];
    block_sum<NR>(pq, tot);
    if (threadIdx.x == 0)
        for (int c = 0; c < NR; ++c) a.tpart[t * NR + c] = tot[c];
}
template <int NR, int G>
__global__ __launch_bounds__(BT) void k_bspmm(BArgs a) {
    bspmm_tile<NR, G>(a, blockIdx.x);
}

// per (system, right-hand side) pair, one wavefront: pq = sum of the
// system's tile partials (lane-strided, then a fixed butterfly), alpha = rz / pq
// (k: the pair, one whole wavefront per call)
template <int NR>
__device__ inline void breduce_pq_pair(const BArgs& a, int k) {
    const int lane = threadIdx.x & 63;
    const int s = k / NR, c = k % NR;
    if (a.done[k]) return;
    double pq = 0.0;
    for (int64_t t = a.sys_tile[s] + lane; t < a.sys_tile[s + 1]; t += 64) pq += a.tpart[t * NR + c];
    for (int m = 32; m >= 1; m >>= 1) pq += __shfl_xor(pq, m, 64);
    if (lane == 0) {
        double* sc = a.scal + (size_t)k * S_N;
        sc[S_PQ] = pq;
        sc[S_ALPHA] = pq != 0.0 ? sc[S_RZ] / pq : 0.0;
    }
}
template <int NR>
__global__ __launch_bounds__(256) void k_breduce_pq(BArgs a) {
    const int k = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64);
    if (k >= a.n_sys * NR) return;  // uniform per wavefront
    breduce_pq_pair<NR>(a, k);
}

// x += alpha p, r -= alpha q, z = dinv r; partial r.z and r.r
template <int NR>
__device__ inline void bupdate_tile(const BArgs& a, int64_t li) {
    const int64_t t = a.tile_list[li];
    const int s = a.tile_sys[t];
    if (tile_done<NR>(a, s)) return;
    double alpha[NR];
    bool act[NR];
    for (int c = 0; c < NR; ++c) {
        alpha[c] = a.scal[(s * NR + c) * S_N + S_ALPHA];
        act[c] = a.done[s * NR + c] == 0;
    }
    double rz[NR], rr[NR];
    for (int c = 0; c < NR; ++c) rz[c] = rr[c] = 0.0;
    for (int64_t r = a.tile_row[t] + threadIdx.x; r < a.tile_row[t + 1]; r += BT) {
        const double di = a.dinv[r];
        for (int c = 0; c < NR; ++c) {
            if (!act[c]) continue;
            const int64_t i = r * NR + c;
            a.x[i] += alpha[c] * a.pnew[i];
            const double rn = a.r[i] - alpha[c] * a.q[i];
            a.r[i] = rn;
            const double zn = di * rn;
            a.z[i] = zn;
            rz[c] += rn * zn;
            rr[c] += rn * rn;
        }
    }
    double trz[NR], trr[NR];
    block_sum<NR>(rz, trz);
    block_sum<NR>(rr, trr);
    if (threadIdx.x == 0)
        for (int c = 0; c < NR; ++c) {
            a.tpart2[t * 2 * NR + c] = trz[c];
            a.tpart2[t * 2 * NR + NR + c] = trr[c];
        }
}
template <int NR>
__global__ __launch_bounds__(BT) void k_bupdate(BArgs a) {
    bupdate_tile<NR>(a, blockIdx.x);
}

// per pair, one wavefront: beta = rz_new / rz, convergence test rr <= tol2,
// iteration count; every pair still running afterwards sets *active
template <int NR>
__device__ inline void breduce_rz_pair(const BArgs& a, int k, int maxit, int* active) {
    const int lane = threadIdx.x & 63;
    if (a.done[k]) return;
    const int s = k / NR, c = k % NR;
    double rz = 0.0, rr = 0.0;
    for (int64_t t = a.sys_tile[s] + lane; t < a.sys_tile[s + 1]; t += 64) {
        rz += a.tpart2[t * 2 * NR + c];
        rr += a.tpart2[t * 2 * NR + NR + c];
    }
    for (int m = 32; m >= 1; m >>= 1) {
        rz += __shfl_xor(rz, m, 64);
        rr += __shfl_xor(rr, m, 64);
    }
    if (lane != 0) return;
    double* sc = a.scal + (size_t)k * S_N;
    sc[S_BETA] = sc[S_RZ] != 0.0 ? rz / sc[S_RZ] : 0.0;
    sc[S_RZ] = rz;
    sc[S_RR] = rr;
    sc[S_IT] += 1.0;
    if (rr <= sc[S_TOL2] || sc[S_IT] >= maxit || !(rr == rr)) {
        a.done[k] = 1;
    } else {
        *active = 1;
    }
}
template <int NR>
__global__ __launch_bounds__(256) void k_breduce_rz(BArgs a, int maxit) {
    const int k = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64);
    if (k >= a.n_sys * NR) return;  // uniform per wavefront
    breduce_rz_pair<NR>(a, k, maxit, a.active);
}

// The whole solve in one cooperative launch: every block loops over the
// iterations, takes launch-list entries blockIdx.x, blockIdx.x + gridDim.x, ...
// in each tile phase and pairs wave by wave in each reduction phase, with a
// grid-wide barrier between phases.  Same device functions as the kernels
// above, so the same bits.  The running flag alternates between two slots:
// iteration it raises active[it & 1] and clears active[(it + 1) & 1], which
// every block read for the last time four barriers earlier.
template <int NR, int G>
__global__ __launch_bounds__(BT) void k_bpersist(BArgs a, int maxit, int* active) {
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const int waves = (int)(gridDim.x * (BT / 64));
    const int wave = (int)(blockIdx.x * (BT / 64) + threadIdx.x / 64);
    const int pairs = a.n_sys * NR;
    for (int it = 0; it <= maxit; ++it) {
        for (int64_t li = blockIdx.x; li < a.n_launch; li += gridDim.x) bpupd_tile<NR>(a, li);
        grid.sync();
        for (int64_t li = blockIdx.x; li < a.n_launch; li += gridDim.x) bspmm_tile<NR, G>(a, li);
        grid.sync();
        for (int k = wave; k < pairs; k += waves) breduce_pq_pair<NR>(a, k);
        grid.sync();
        for (int64_t li = blockIdx.x; li < a.n_launch; li += gridDim.x) bupdate_tile<NR>(a, li);
        grid.sync();
        if (blockIdx.x == 0 && threadIdx.x == 0) active[(it + 1) & 1] = 0;
        for (int k = wave; k < pairs; k += waves) breduce_rz_pair<NR>(a, k, maxit, &active[it & 1]);
        grid.sync();
        // one coherent read per block (an atomic at L2, not an L1 line that
        // another block's earlier value may still occupy), shared with the block
        __shared__ int run;
        if (threadIdx.x == 0) run = atomicAdd(&active[it & 1], 0);
        __syncthreads();
        if (run == 0) break;  // the same value in every block
        __syncthreads();      // every thread has read run before thread 0 rewrites it
    }
}

__global__ void k_bclear_active(int* active) { *active = 0; }

// set-up: r = b - K x0 (x0 = 0 here: r = b), z = dinv r, p = 0, beta = 0,
// rz, rr0, tol2 per pair
template <int NR>
__global__ __launch_bounds__(BT) void k_binit(BArgs a) {
    const int64_t t = blockIdx.x;
    double rz[NR], rr[NR];
    for (int c = 0; c < NR; ++c) rz[c] = rr[c] = 0.0;
    for (int64_t r = a.tile_row[t] + threadIdx.x; r < a.tile_row[t + 1]; r += BT) {
        const double di = a.dinv[r];
        for (int c = 0; c < NR; ++c) {
            const int64_t i = r * NR + c;
            const double rn = a.r[i];
            const double zn = di * rn;
            a.x[i] = 0.0;
            a.z[i] = zn;
            a.pnew[i] = 0.0;
            rz[c] += rn * zn;
            rr[c] += rn * rn;
        }
    }
    double trz[NR], trr[NR];
    block_sum<NR>(rz, trz);
    block_sum<NR>(rr, trr);
    if (threadIdx.x == 0)
        for (int c = 0; c < NR; ++c) {
            a.tpart2[t * 2 * NR + c] = trz[c];
            a.tpart2[t * 2 * NR + NR + c] = trr[c];
        }
}

template <int NR>
__global__ void k_binit_scal(BArgs a, double rtol) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n_sys * NR) return;
    const int s = k / NR, c = k % NR;
    double rz = 0.0, rr = 0.0;
    for (int64_t t = a.sys_tile[s]; t < a.sys_tile[s + 1]; ++t) {
        rz += a.tpart2[t * 2 * NR + c];
        rr += a.tpart2[t * 2 * NR + NR + c];
    }
    double* sc = a.scal + (size_t)k * S_N;
    sc[S_RZ] = rz;
    sc[S_RR] = rr;
    sc[S_RR0] = rr;
    sc[S_TOL2] = rtol * rtol * rr;
    sc[S_BETA] = 0.0;
    sc[S_ALPHA] = 0.0;
    sc[S_IT] = 0.0;
    a.done[k] = rr == 0.0 ? 1 : 0;
}

__global__ void k_bdiag(int64_t n, const int32_t* rowptr, const int32_t* rowsplit, const int32_t* colidx, const double* K,
                        double* dinv) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double d = 0.0;
    for (int32_t j = rowptr[r]; j < rowptr[r] + rowsplit[r]; ++j)
        if (colidx[j] == r) d = K[j];
    dinv[r] = d != 0.0 ? 1.0 / d : 0.0;
}

// b[c][r] = -sum_{prescribed cols j of row r} K[j] up[c][col - n_free], free rows
__global__ void k_rhs_pres(int64_t nf, int nrhs, int64_t np, const int32_t* rowptr, const int32_t* rowsplit,
                           const int32_t* colidx, const double* K, const double* up, double* b) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nf) return;
    for (int c = 0; c < nrhs; ++c) {
        double s = 0.0;
        for (int32_t j = rowptr[r] + rowsplit[r]; j < rowptr[r + 1]; ++j) s += K[j] * up[(size_t)c * np + (colidx[j] - nf)];
        b[(size_t)c * nf + r] = -s;
    }
}

// reactions on the prescribed rows: R[c][p] = sum_j K[j] u_c[col] (free and prescribed columns)
__global__ void k_reactions(int64_t nf, int nrhs, int64_t np, const int32_t* rowptr, const int32_t* colidx, const double* K,
                            const double* uf, const double* up, double* R) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int64_t r = nf + p;
    for (int c = 0; c < nrhs; ++c) {
        double s = 0.0;
        for (int32_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
            const int64_t col = colidx[j];
            s += K[j] * (col < nf ? uf[(size_t)c * nf + col] : up[(size_t)c * np + (col - nf)]);
        }
        R[(size_t)c * np + p] = s;
    }
}

}  // namespace pgfb

// ===========================================================================
using namespace pgfb;

#define BCHK(x)                                                                          \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) throw std::string("HIP error: ") + hipGetErrorString(e_); \
    } while (0)

namespace {

template <typename T>
T* balloc(size_t n) {
    void* p = nullptr;
    BCHK(hipMalloc(&p, sizeof(T) * (n ? n : 1)));
    return (T*)p;
}

struct Tiles {
    std::vector<int32_t> sys;
    std::vector<int64_t> row, sys_tile;
};

Tiles make_tiles(const std::vector<int64_t>& sys_ptr, int64_t rows_per_tile) {
    Tiles T;
    const int n_sys = (int)sys_ptr.size() - 1;
    T.row.push_back(sys_ptr[0]);
    T.sys_tile.push_back(0);
    for (int s = 0; s < n_sys; ++s) {
        for (int64_t r = sys_ptr[s]; r < sys_ptr[s + 1]; r += rows_per_tile) {
            T.sys.push_back(s);
            T.row.push_back(std::min(r + rows_per_tile, sys_ptr[s + 1]));
        }
        T.sys_tile.push_back((int64_t)T.sys.size());
    }
    return T;
}

struct BatchRun {
    BArgs a{};
    int32_t *d_tile_sys = nullptr, *d_tile_list = nullptr;
    int64_t *d_tile_row = nullptr, *d_sys_tile = nullptr;
    ~BatchRun() {
        void* p[] = {d_tile_sys, d_tile_list, d_tile_row, d_sys_tile, a.x, a.r, a.z, a.pnew, a.q, a.tpart,
                     a.tpart2, a.scal, a.done, a.active, (void*)a.dinv};
        for (void* q : p)
            if (q) (void)hipFree(q);
    }
};

template <int NR, int G>
void enqueue_iteration(BArgs& a, hipStream_t st, int maxit) {
    const unsigned nt = (unsigned)a.n_launch;
    const unsigned np = (unsigned)(((int64_t)a.n_sys * NR * 64 + 255) / 256);  // one wavefront per pair
    hipLaunchKernelGGL(k_bpupd<NR>, dim3(nt), dim3(BT), 0, st, a);
    hipLaunchKernelGGL((k_bspmm<NR, G>), dim3(nt), dim3(BT), 0, st, a);
    hipLaunchKernelGGL(k_breduce_pq<NR>, dim3(np), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_bupdate<NR>, dim3(nt), dim3(BT), 0, st, a);
    hipLaunchKernelGGL(k_breduce_rz<NR>, dim3(np), dim3(256), 0, st, a, maxit);
}

template <int NR>
void enqueue_iteration_g(int G, BArgs& a, hipStream_t st, int maxit) {
    if (G == 4) enqueue_iteration<NR, 4>(a, st, maxit);
    else if (G == 8) enqueue_iteration<NR, 8>(a, st, maxit);
    else enqueue_iteration<NR, 16>(a, st, maxit);
}

constexpr int64_t PERSIST_MAX_TILES = 256;

template <int NR>
const void* persist_kernel(int G) {
    if (G == 4) return reinterpret_cast<const void*>(&k_bpersist<NR, 4>);
    if (G == 8) return reinterpret_cast<const void*>(&k_bpersist<NR, 8>);
    return reinterpret_cast<const void*>(&k_bpersist<NR, 16>);
}

template <int NR>
int run_batch(pgfd::Ctx& c, const std::vector<int64_t>& sys_ptr, const double* b, double* x, double rtol, int maxit,
              int32_t* iters, double* relres, double* ms_out, int32_t* tuned, int G_req, int rows_req,
              bool persistent) {
    const int n_sys = (int)sys_ptr.size() - 1;
    const int64_t nf = c.n_free;
    // autotuner candidates: (threads per row, rows per tile)
    std::vector<std::pair<int, int>> cand;
    if (G_req > 0 && rows_req > 0) cand.push_back({G_req, rows_req});
    else
        for (int G : {4, 8, 16})
            for (int R : {128, 512, 2048}) cand.push_back({G, R});
    BatchRun run;
    BArgs& a = run.a;
    a.n_sys = n_sys;
    a.rowptr = c.rowptr;
    a.rowsplit = c.rowsplit;
    a.colidx = c.colidx;
    a.K = c.K;
    double* dinv = balloc<double>(nf);
    a.dinv = dinv;
    hipLaunchKernelGGL(k_bdiag, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, c.stream, nf, (const int32_t*)c.rowptr,
                       (const int32_t*)c.rowsplit, (const int32_t*)c.colidx, (const double*)c.K, dinv);
    a.x = balloc<double>(nf * NR);
    a.r = balloc<double>(nf * NR);
    a.z = balloc<double>(nf * NR);
    a.pnew = balloc<double>(nf * NR);  // p, updated in place
    a.q = balloc<double>(nf * NR);
    a.scal = balloc<double>((size_t)n_sys * NR * S_N);
    a.done = balloc<int>((size_t)n_sys * NR);
    a.active = balloc<int>(1);
    int* h_active = nullptr;
    BCHK(hipHostMalloc((void**)&h_active, sizeof(int), hipHostMallocDefault));
    // b interleaved [nf][NR] from [NR][nf]
    std::vector<double> bi((size_t)nf * NR);
    for (int cc = 0; cc < NR; ++cc)
        for (int64_t r = 0; r < nf; ++r) bi[r * NR + cc] = b[(size_t)cc * nf + r];
    std::vector<double> r_host = bi;
    Tiles T;
    auto set_tiles = [&](int rows) {
        T = make_tiles(sys_ptr, rows);
        if (run.d_tile_sys) {
            (void)hipFree(run.d_tile_sys); (void)hipFree(run.d_tile_row); (void)hipFree(run.d_sys_tile);
            (void)hipFree(run.d_tile_list);
        }
        if (a.tpart) { (void)hipFree(a.tpart); (void)hipFree(a.tpart2); }
        run.d_tile_sys = balloc<int32_t>(T.sys.size());
        run.d_tile_row = balloc<int64_t>(T.row.size());
        run.d_sys_tile = balloc<int64_t>(T.sys_tile.size());
        BCHK(hipMemcpy(run.d_tile_sys, T.sys.data(), sizeof(int32_t) * T.sys.size(), hipMemcpyHostToDevice));
        BCHK(hipMemcpy(run.d_tile_row, T.row.data(), sizeof(int64_t) * T.row.size(), hipMemcpyHostToDevice));
        BCHK(hipMemcpy(run.d_sys_tile, T.sys_tile.data(), sizeof(int64_t) * T.sys_tile.size(), hipMemcpyHostToDevice));
        a.n_tiles = (int64_t)T.sys.size();
        std::vector<int32_t> ident(T.sys.size());
        for (size_t k = 0; k < ident.size(); ++k) ident[k] = (int32_t)k;
        run.d_tile_list = balloc<int32_t>(ident.size());
        BCHK(hipMemcpy(run.d_tile_list, ident.data(), sizeof(int32_t) * ident.size(), hipMemcpyHostToDevice));
        a.tile_list = run.d_tile_list;
        a.n_launch = a.n_tiles;
        a.tile_sys = run.d_tile_sys;
        a.tile_row = run.d_tile_row;
        a.sys_tile = run.d_sys_tile;
        a.tpart = balloc<double>(a.n_tiles * NR);
        a.tpart2 = balloc<double>(a.n_tiles * 2 * NR);
    };
    auto init = [&]() {
        BCHK(hipMemcpyAsync(a.r, bi.data(), sizeof(double) * bi.size(), hipMemcpyHostToDevice, c.stream));
        hipLaunchKernelGGL(k_binit<NR>, dim3((unsigned)a.n_tiles), dim3(BT), 0, c.stream, a);
        hipLaunchKernelGGL(k_binit_scal<NR>, dim3((unsigned)((n_sys * NR + 255) / 256)), dim3(256), 0, c.stream, a, rtol);
        BCHK(hipGetLastError());
    };
    // ---- autotune: a few iterations of every candidate from the same start ----
    int bestG = cand[0].first, bestR = cand[0].second;
    if (cand.size() > 1) {
        double best = 1e300;
        hipEvent_t e0, e1;
        BCHK(hipEventCreate(&e0));
        BCHK(hipEventCreate(&e1));
        for (auto [G, R] : cand) {
            set_tiles(R);
            init();
            enqueue_iteration_g<NR>(G, a, c.stream, maxit);  // warm-up
            BCHK(hipEventRecord(e0, c.stream));
            for (int k = 0; k < 4; ++k) enqueue_iteration_g<NR>(G, a, c.stream, maxit);
            BCHK(hipEventRecord(e1, c.stream));
            BCHK(hipEventSynchronize(e1));
            float ms;
            BCHK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; bestG = G; bestR = R; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    if (tuned) { tuned[0] = bestG; tuned[1] = bestR; }
    set_tiles(bestR);
    init();
    if (persistent) {
        // ---- the whole solve in one cooperative launch ----
        int* act2 = balloc<int>(2);
        BCHK(hipMemsetAsync(act2, 0, 2 * sizeof(int), c.stream));
        int dev = 0, per_cu = 0;
        BCHK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        BCHK(hipGetDeviceProperties(&prop, dev));
        const void* kfn = persist_kernel<NR>(bestG);
        BCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, BT, 0));
        // one block per tile, at most one per CU: the size this path has been
        // run and checked at; larger batches are refused (their runs did not
        // finish within minutes where the graph replays took 15.7 s)
        if (a.n_launch > PERSIST_MAX_TILES || a.n_launch > (int64_t)prop.multiProcessorCount) {
            (void)hipFree(act2);
            (void)hipHostFree(h_active);
            throw std::string("persistent batch launch holds at most ") + std::to_string(PERSIST_MAX_TILES) +
                " row tiles (one per CU); this batch has " + std::to_string(a.n_launch) + ": use pgf_batch_pcg";
        }
        const int64_t grid = std::min<int64_t>((int64_t)per_cu * prop.multiProcessorCount, std::max<int64_t>(a.n_launch, 1));
        if (grid < 1) {
            (void)hipFree(act2);
            (void)hipHostFree(h_active);
            throw std::string("the persistent batch kernel cannot be resident on this device");
        }
        int mi = maxit;
        void* args[] = {(void*)&a, (void*)&mi, (void*)&act2};
        hipEvent_t t0, t1;
        BCHK(hipEventCreate(&t0));
        BCHK(hipEventCreate(&t1));
        BCHK(hipEventRecord(t0, c.stream));
        BCHK(hipLaunchCooperativeKernel(kfn, dim3((unsigned)grid), dim3(BT), args, 0, c.stream));
        BCHK(hipEventRecord(t1, c.stream));
        BCHK(hipEventSynchronize(t1));
        float ms;
        BCHK(hipEventElapsedTime(&ms, t0, t1));
        if (ms_out) *ms_out = ms;
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        (void)hipFree(act2);
    } else {
    // ---- capture CHUNK iterations in one graph, replay until done ----
    const int CHUNK = 16;
    int* h_done = nullptr;
    BCHK(hipHostMalloc((void**)&h_done, sizeof(int) * (size_t)n_sys * NR, hipHostMallocDefault));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto capture = [&]() {
        if (exec) { (void)hipGraphExecDestroy(exec); (void)hipGraphDestroy(graph); }
        BArgs a0 = a;
        BCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        for (int k = 0; k < CHUNK; ++k) {
            // the running flag is taken in the chunk's last iteration
            if (k == CHUNK - 1) hipLaunchKernelGGL(k_bclear_active, dim3(1), dim3(1), 0, c.stream, a.active);
            enqueue_iteration_g<NR>(bestG, a, c.stream, maxit);
        }
        BCHK(hipMemcpyAsync(h_active, a.active, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        BCHK(hipMemcpyAsync(h_done, a.done, sizeof(int) * (size_t)n_sys * NR, hipMemcpyDeviceToHost, c.stream));
        BCHK(hipStreamEndCapture(c.stream, &graph));
        BCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        a = a0;
    };
    capture();
    hipEvent_t t0, t1;
    BCHK(hipEventCreate(&t0));
    BCHK(hipEventCreate(&t1));
    BCHK(hipEventRecord(t0, c.stream));
    int chunks = 0;
    std::vector<int32_t> live;
    for (;;) {
        BCHK(hipGraphLaunch(exec, c.stream));
        BCHK(hipStreamSynchronize(c.stream));
        ++chunks;
        if (*h_active == 0 || chunks * CHUNK >= maxit + CHUNK) break;
        // compact the launch list to the tiles of systems with a running pair
        live.clear();
        for (int s = 0; s < n_sys; ++s) {
            bool run_s = false;
            for (int cc = 0; cc < NR; ++cc) run_s = run_s || h_done[s * NR + cc] == 0;
            if (run_s)
                for (int64_t t = T.sys_tile[s]; t < T.sys_tile[s + 1]; ++t) live.push_back((int32_t)t);
        }
        if (!live.empty() && (double)live.size() < 0.6 * (double)a.n_launch) {
            BCHK(hipMemcpyAsync(run.d_tile_list, live.data(), sizeof(int32_t) * live.size(), hipMemcpyHostToDevice,
                                c.stream));
            BCHK(hipStreamSynchronize(c.stream));
            a.n_launch = (int64_t)live.size();
            capture();
        }
    }
    BCHK(hipEventRecord(t1, c.stream));
    BCHK(hipEventSynchronize(t1));
    float ms;
    BCHK(hipEventElapsedTime(&ms, t0, t1));
    if (ms_out) *ms_out = ms;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    (void)hipHostFree(h_done);
    }
    // results
    std::vector<double> xi((size_t)nf * NR), sc((size_t)n_sys * NR * S_N);
    BCHK(hipMemcpy(xi.data(), a.x, sizeof(double) * xi.size(), hipMemcpyDeviceToHost));
    BCHK(hipMemcpy(sc.data(), a.scal, sizeof(double) * sc.size(), hipMemcpyDeviceToHost));
    for (int cc = 0; cc < NR; ++cc)
        for (int64_t r = 0; r < nf; ++r) x[(size_t)cc * nf + r] = xi[r * NR + cc];
    int all_conv = 1;
    for (int k = 0; k < n_sys * NR; ++k) {
        const double* s = &sc[(size_t)k * S_N];
        if (iters) iters[k] = (int32_t)s[S_IT];
        if (relres) relres[k] = s[S_RR0] > 0 ? std::sqrt(s[S_RR] / s[S_RR0]) : 0.0;
        if (!(s[S_RR] <= s[S_TOL2])) all_conv = 0;
    }
    (void)hipHostFree(h_active);
    return all_conv ? 0 : PGF_ENOTCONVERGED;
}

}  // namespace

extern "C" {

// b = -K_fp u_p for nrhs prescribed-value vectors (host [nrhs][n_pres] -> host [nrhs][n_free])
PGF_API int pgf_rhs_prescribed(pgf_ctx* h, int32_t nrhs, const double* up, double* b) {
    if (!h || nrhs < 1 || !up || !b) return PGF_EARG;
    pgfd::Ctx* c = pgfd::ctx_of(h);
    try {
        const int64_t nf = c->n_free, np = c->n_pres;
        double* dup = balloc<double>((size_t)nrhs * np);
        double* db = balloc<double>((size_t)nrhs * nf);
        BCHK(hipMemcpy(dup, up, sizeof(double) * nrhs * np, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_rhs_pres, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, c->stream, nf, nrhs, np,
                           (const int32_t*)c->rowptr, (const int32_t*)c->rowsplit, (const int32_t*)c->colidx,
                           (const double*)c->K, (const double*)dup, db);
        BCHK(hipGetLastError());
        BCHK(hipStreamSynchronize(c->stream));
        BCHK(hipMemcpy(b, db, sizeof(double) * nrhs * nf, hipMemcpyDeviceToHost));
        (void)hipFree(dup);
        (void)hipFree(db);
        return 0;
    } catch (const std::string& e) {
        c->err = e;
        return PGF_EDEVICE;
    }
}

// reactions R = [K_pf K_pp] [u_f; u_p] (needs an assembly with PGF_ASM_ALL_ROWS);
// host uf [nrhs][n_free], up [nrhs][n_pres] -> R [nrhs][n_pres]
PGF_API int pgf_reactions(pgf_ctx* h, int32_t nrhs, const double* uf, const double* up, double* R) {
    if (!h || nrhs < 1 || !uf || !up || !R) return PGF_EARG;
    pgfd::Ctx* c = pgfd::ctx_of(h);
    try {
        const int64_t nf = c->n_free, np = c->n_pres;
        double* duf = balloc<double>((size_t)nrhs * nf);
        double* dup = balloc<double>((size_t)nrhs * np);
        double* dR = balloc<double>((size_t)nrhs * np);
        BCHK(hipMemcpy(duf, uf, sizeof(double) * nrhs * nf, hipMemcpyHostToDevice));
        BCHK(hipMemcpy(dup, up, sizeof(double) * nrhs * np, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_reactions, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, nf, nrhs, np,
                           (const int32_t*)c->rowptr, (const int32_t*)c->colidx, (const double*)c->K, (const double*)duf,
                           (const double*)dup, dR);
        BCHK(hipGetLastError());
        BCHK(hipStreamSynchronize(c->stream));
        BCHK(hipMemcpy(R, dR, sizeof(double) * nrhs * np, hipMemcpyDeviceToHost));
        (void)hipFree(duf);
        (void)hipFree(dup);
        (void)hipFree(dR);
        return 0;
    } catch (const std::string& e) {
        c->err = e;
        return PGF_EDEVICE;
    }
}

// Solve K_ff X = B for a block-diagonal K_ff (blocks = systems given by their
// free-row ranges sys_ptr[0..n_sys]), nrhs right-hand sides per system, zero
// start, each (system, rhs) pair to ||r|| <= rtol ||b||.  b and x are host
// arrays [nrhs][n_free].  tune[0..1] in: threads per row and rows per tile
// (0: autotune); out: the choice used.
static int batch_pcg(pgf_ctx* h, int32_t n_sys, const int64_t* sys_ptr, int32_t nrhs, const double* b, double* x,
                     double rtol, int32_t maxit, int32_t* iters, double* relres, double* ms, int32_t* tune,
                     bool persistent) {
    if (!h || n_sys <= 0 || !sys_ptr || !b || !x || nrhs < 1 || nrhs > 4 || maxit < 1) return PGF_EARG;
    pgfd::Ctx* c = pgfd::ctx_of(h);
    try {
        std::vector<int64_t> sp(sys_ptr, sys_ptr + n_sys + 1);
        if (sp[0] != 0 || sp[n_sys] != c->n_free) throw std::string("system ranges must cover the free rows");
        for (int s = 0; s < n_sys; ++s)
            if (sp[s + 1] < sp[s]) throw std::string("system ranges must be ascending");
        // the free-free block must be block diagonal over these ranges
        int s = 0;
        for (int64_t r = 0; r < c->n_free; ++r) {
            while (r >= sp[s + 1]) ++s;
            for (int64_t j = c->h_rowptr[r]; j < c->h_rowptr[r] + c->h_rowsplit[r]; ++j)
                if (c->h_colidx[j] < sp[s] || c->h_colidx[j] >= sp[s + 1])
                    throw std::string("K_ff couples two systems: not block diagonal over the given ranges");
        }
        const int Greq = tune ? tune[0] : 0, Rreq = tune ? tune[1] : 0;
        if (Greq && Greq != 4 && Greq != 8 && Greq != 16) throw std::string("threads per row must be 4, 8 or 16");
        int rc;
        const bool pe = persistent;
        if (nrhs == 1) rc = run_batch<1>(*c, sp, b, x, rtol, maxit, iters, relres, ms, tune, Greq, Rreq, pe);
        else if (nrhs == 2) rc = run_batch<2>(*c, sp, b, x, rtol, maxit, iters, relres, ms, tune, Greq, Rreq, pe);
        else if (nrhs == 3) rc = run_batch<3>(*c, sp, b, x, rtol, maxit, iters, relres, ms, tune, Greq, Rreq, pe);
        else rc = run_batch<4>(*c, sp, b, x, rtol, maxit, iters, relres, ms, tune, Greq, Rreq, pe);
        return rc;
    } catch (const std::string& e) {
        c->err = e;
        if (e.rfind("persistent", 0) == 0) return PGF_EUNSUPPORTED;
        return e.rfind("HIP", 0) == 0 ? PGF_EDEVICE : PGF_EARG;
    }
}

PGF_API int pgf_batch_pcg(pgf_ctx* h, int32_t n_sys, const int64_t* sys_ptr, int32_t nrhs, const double* b, double* x,
                          double rtol, int32_t maxit, int32_t* iters, double* relres, double* ms, int32_t* tune) {
    return batch_pcg(h, n_sys, sys_ptr, nrhs, b, x, rtol, maxit, iters, relres, ms, tune, false);
}

PGF_API int pgf_batch_pcg_persistent(pgf_ctx* h, int32_t n_sys, const int64_t* sys_ptr, int32_t nrhs, const double* b,
                                     double* x, double rtol, int32_t maxit, int32_t* iters, double* relres, double* ms,
                                     int32_t* tune) {
    return batch_pcg(h, n_sys, sys_ptr, nrhs, b, x, rtol, maxit, iters, relres, ms, tune, true);
}

}  // extern "C"
