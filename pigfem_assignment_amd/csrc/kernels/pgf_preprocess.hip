// IGFEM preprocessing on the device: node detection, element detection,
// enrichment nodes and the integration tables of the cut elements, the work
// of Problem::init's analyze_nodes, analyze_elements, add_enrichments and
// store_shape_functions (src/Mesh.cpp:2676-3202, 1443-1605).
//
// Every per-node, per-element and per-edge step runs the same source as the
// host path (host/pgf_igfem.hpp) in one thread; this file is built without
// multiply-add contraction, so the flags, the enrichment coordinates and the
// tables equal the host's bit for bit.  Layout: one thread per node, per
// element, per unique cut edge or per cut element, grid-stride loops.
//
//   nodes     k_pp_nodes: detection value per node.  A node within 1e-12 of a
//             surface is flagged; those (rare) nodes are moved and detected
//             on the host in node order, as the host path does, because a
//             move reads the coordinates of a neighbour that may have moved.
//   elements  k_pp_elem_flags (cut flag, material of uncut elements), an
//             exclusive scan of the flags gives every cut element its index in
//             element order, k_pp_cut_fill writes the cut-element records.
//   enrich    k_pp_keys packs each cut edge as (min node, max node, inclusion)
//             in 64 bits; radix sort + unique give the enrichment nodes in the
//             host's sorted order; k_pp_roots places each on its edge
//             (Steffenson / Brent), k_pp_enrich_ids binary-searches every
//             cut edge's node.
//   tables    k_pp_tab_sizes, three scans for the offsets, k_pp_tables.
//
// Errors found on the device (an element the tables do not hold, an edge with
// no root, a bad inverse map) re-run that phase on the host, which raises the
// host path's error.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <climits>
#include <string>
#include <vector>

#include "../host/pgf_igfem.hpp"
#include "../host/pgf_mesh.hpp"

static_assert(sizeof(pgf::CutElem) == 224, "cut-element record layout differs between host and device");

namespace pgf {

namespace {

#define PPCHK(x)                                                                                     \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess)                                                                        \
            throw Error(PGF_ERR_DEVICE, std::string("device set-up: ") + #x + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        if (count) PPCHK(hipMalloc(&p, sizeof(T) * count));
    }
    void up(const T* h, size_t count) {
        alloc(count);
        if (count) PPCHK(hipMemcpy(p, h, sizeof(T) * count, hipMemcpyHostToDevice));
    }
    void down(T* h, size_t count) const {
        if (count) PPCHK(hipMemcpy(h, p, sizeof(T) * count, hipMemcpyDeviceToHost));
    }
};

constexpr int PP_THREADS = 256;

unsigned pp_grid(int64_t n) {
    int64_t b = (n + PP_THREADS - 1) / PP_THREADS;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return (unsigned)b;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// packed cut-edge key: (min node, max node, inclusion) in 26 + 26 + 12 bits,
// ordered as the host's lexicographic (min, max, inclusion) keys
constexpr int KEY_NODE_BITS = 26, KEY_INCL_BITS = 12;
constexpr uint64_t KEY_NONE = ~0ull;

__device__ inline uint64_t pack_key(int32_t g0, int32_t g1, int32_t inc) {
    const uint64_t lo = (uint64_t)(g0 < g1 ? g0 : g1), hi = (uint64_t)(g0 < g1 ? g1 : g0);
    return (lo << (KEY_NODE_BITS + KEY_INCL_BITS)) | (hi << KEY_INCL_BITS) | (uint64_t)inc;
}

__global__ void k_pp_nodes(int64_t n, const double* __restrict__ X0, const double* __restrict__ X1,
                           const double* __restrict__ X2, const Inclusion* __restrict__ incl, int ni, int zlim,
                           int32_t* __restrict__ flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = igfem::scan_node(incl, ni, zlim != 0, X0[i], X1[i], X2[i]);
}

__device__ inline int classify(int kind, const int* nd, int emat, const Inclusion* incl, bool coh, CutElem& ce,
                               bool& cut) {
    return kind == QUAD4 ? igfem::classify_quad4(nd, emat, incl, coh, ce, cut)
                         : igfem::classify_tet4(nd, emat, incl, coh, ce, cut);
}

__global__ void k_pp_elem_flags(int64_t ne, int kind, const int32_t* __restrict__ conn,
                                const int32_t* __restrict__ flag, int32_t* __restrict__ emat,
                                const Inclusion* __restrict__ incl, int coh, int32_t* __restrict__ cutf,
                                int* __restrict__ err_elem) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        int nd[4];
        for (int k = 0; k < 4; ++k) nd[k] = flag[conn[e * 4 + k]];
        CutElem ce;
        bool cut = false;
        const int st = classify(kind, nd, emat[e], incl, coh != 0, ce, cut);
        if (st != igfem::CLS_OK) {
            atomicMin(err_elem, (int)e);
            cutf[e] = 0;
            continue;
        }
        cutf[e] = cut ? 1 : 0;
        if (!cut && nd[0] != -1) emat[e] = incl[nd[0]].mat;
    }
}

__global__ void k_pp_cut_fill(int64_t ne, int kind, const int32_t* __restrict__ conn, const int32_t* __restrict__ flag,
                              const int32_t* __restrict__ emat, const Inclusion* __restrict__ incl, int coh,
                              int default_law, const int32_t* __restrict__ cutf, const int32_t* __restrict__ pos,
                              CutElem* __restrict__ cut, int32_t* __restrict__ cut_index) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        if (!cutf[e]) {
            cut_index[e] = -1;
            continue;
        }
        int nd[4];
        for (int k = 0; k < 4; ++k) nd[k] = flag[conn[e * 4 + k]];
        CutElem ce;
        bool is_cut = false;
        classify(kind, nd, emat[e], incl, coh != 0, ce, is_cut);
        ce.elem = (int32_t)e;
        igfem::cohesive_laws(ce, incl, default_law);
        cut[pos[e]] = ce;
        cut_index[e] = pos[e];
    }
}

__global__ void k_pp_keys(int64_t nc, int kind, const int32_t* __restrict__ conn, const CutElem* __restrict__ cut,
                          uint64_t* __restrict__ keys) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (int64_t)gridDim.x * blockDim.x) {
        const CutElem& ce = cut[i];
        for (int k = 0; k < 4; ++k) {
            uint64_t key = KEY_NONE;
            if (k < ce.n_cut) {
                int a, b;
                igfem::edge_nodes(kind, ce.cut_edge[k], a, b);
                key = pack_key(conn[(int64_t)ce.elem * 4 + a], conn[(int64_t)ce.elem * 4 + b], ce.enrich_incl[k]);
            }
            keys[i * 4 + k] = key;
        }
    }
}

__global__ void k_pp_roots(int64_t n_init, const uint64_t* __restrict__ uniq, int64_t n_nodes, int mirror,
                           const Inclusion* __restrict__ incl, double* __restrict__ X0, double* __restrict__ X1,
                           double* __restrict__ X2, int* __restrict__ err) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_init; k += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = uniq[k];
        const int64_t g0 = (int64_t)(key >> (KEY_NODE_BITS + KEY_INCL_BITS));
        const int64_t g1 = (int64_t)((key >> KEY_INCL_BITS) & ((1ull << KEY_NODE_BITS) - 1));
        const int inc = (int)(key & ((1ull << KEY_INCL_BITS) - 1));
        const double p0[3] = {X0[g0], X1[g0], X2[g0]}, p1[3] = {X0[g1], X1[g1], X2[g1]};
        double x[3];
        if (!igfem::enrichment_point(incl[inc], p0, p1, x)) {
            atomicMin(err, (int)k);
            continue;
        }
        X0[n_nodes + k] = x[0];
        X1[n_nodes + k] = x[1];
        X2[n_nodes + k] = x[2];
        if (mirror) {
            X0[n_nodes + n_init + k] = x[0];
            X1[n_nodes + n_init + k] = x[1];
            X2[n_nodes + n_init + k] = x[2];
        }
    }
}

__global__ void k_pp_enrich_ids(int64_t nc, int kind, const int32_t* __restrict__ conn, CutElem* __restrict__ cut,
                                const uint64_t* __restrict__ uniq, int64_t n_init, int mirror) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (int64_t)gridDim.x * blockDim.x) {
        CutElem& ce = cut[i];
        for (int k = 0; k < ce.n_cut; ++k) {
            int a, b;
            igfem::edge_nodes(kind, ce.cut_edge[k], a, b);
            const uint64_t key =
                pack_key(conn[(int64_t)ce.elem * 4 + a], conn[(int64_t)ce.elem * 4 + b], ce.enrich_incl[k]);
            int64_t lo = 0, hi = n_init;  // lower_bound
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (uniq[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            ce.enrich[k] = (int32_t)lo;
            if (mirror) ce.enrich[k + ce.n_cut] = (int32_t)(lo + n_init);
        }
    }
}

__global__ void k_pp_tab_sizes(int64_t nc, int dim, const CutElem* __restrict__ cut, int64_t* __restrict__ vcnt,
                               int64_t* __restrict__ ccnt, int64_t* __restrict__ gcnt) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (int64_t)gridDim.x * blockDim.x) {
        int nv, ncp;
        igfem::table_sizes(dim, cut[i], nv, ncp);
        vcnt[i] = nv;
        ccnt[i] = ncp;
        gcnt[i] = (int64_t)nv * (4 + cut[i].n_enrich) * dim;
    }
}

struct TabArgs {
    int64_t nc, n_nodes;
    int dim, kind;
    double tet_weight_scale;
    const CutElem* cut;
    const int32_t* conn;
    const double *X0, *X1, *X2;
    const geom::PolygonTable* poly;
    const int64_t *vptr, *cptr, *gptr;
    double* vw;
    int32_t* vmat;
    int64_t* vgoff;
    double* vgrad;
    double *cw, *cN, *cR;
    uint8_t *cnodes, *cnh;
    int32_t* claw;
    int* err;
};

__global__ void k_pp_tables(TabArgs a) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.nc; i += (int64_t)gridDim.x * blockDim.x) {
        const CutElem& ce = a.cut[i];
        const int nt = 4 + ce.n_enrich;
        double X[MAX_EL_NODES][3];
        for (int s = 0; s < nt; ++s) {
            const int64_t g = s < 4 ? (int64_t)a.conn[(int64_t)ce.elem * 4 + s] : a.n_nodes + ce.enrich[s - 4];
            X[s][0] = a.X0[g];
            X[s][1] = a.X1[g];
            X[s][2] = a.X2[g];
        }
        igfem::TableOut o;
        const int64_t v0 = a.vptr[i], c0 = a.cptr[i], g0 = a.gptr[i];
        o.vw = a.vw + v0; o.vmat = a.vmat + v0; o.vgoff = a.vgoff + v0; o.vgrad = a.vgrad + g0; o.goff0 = g0;
        o.cw = a.cw + c0; o.cN = a.cN + 4 * c0; o.cR = a.cR + 9 * c0; o.cnodes = a.cnodes + 8 * c0;
        o.cnh = a.cnh + c0; o.claw = a.claw + c0;
        if (igfem::cut_tables(a.dim, a.kind, ce, X, a.tet_weight_scale, *a.poly, o) != igfem::TAB_OK)
            atomicMin(a.err, (int)i);
    }
}

// exclusive prefix sum of cnt[0..n) into off[0..n], off[n] = total
template <typename T>
void exclusive_scan(const T* cnt, T* off, int64_t n) {
    PPCHK(hipMemset(off, 0, sizeof(T)));
    if (n == 0) return;
    size_t bytes = 0;
    PPCHK(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, cnt, off + 1, (int)n));
    DBuf<unsigned char> tmp;
    tmp.alloc(bytes);
    PPCHK(hipcub::DeviceScan::InclusiveSum(tmp.p, bytes, cnt, off + 1, (int)n));
}

int read_err(const DBuf<int>& e) {
    int v = INT_MAX;
    e.down(&v, 1);
    return v;
}

void reset_err(DBuf<int>& e) {
    const int v = INT_MAX;
    e.up(&v, 1);
}

}  // namespace

struct DevicePreprocess {
    static void run(Mesh& m, double* phase_ms);
};

void DevicePreprocess::run(Mesh& m, double* phase_ms) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        throw Error(PGF_ERR_DEVICE, "device set-up requested but no HIP device is available");
    const int64_t N = m._n_nodes, NE = m._n_elem;
    const int ni = (int)m._incl.size();
    for (auto& inc : m._incl)
        if ((inc.type == 0) != (m._dim == 2)) throw Error(PGF_ERR_ARG, "inclusion type does not match mesh dimension");
    if (N >= (1ll << KEY_NODE_BITS) || ni >= (1 << KEY_INCL_BITS) || NE > INT_MAX)
        throw Error(PGF_ERR_LIMIT, "device set-up holds at most 2^26 nodes, 4095 inclusions and 2^31 elements");
    const bool coh = m._opt.cohesive;
    const int kind = m._kind, dim = m._dim;
    double t0 = now_ms();

    // ---- nodes ----
    DBuf<double> X[3];
    for (int d = 0; d < 3; ++d) X[d].up(m._X[d].data(), N);
    DBuf<Inclusion> incl;
    incl.up(m._incl.data(), ni);
    DBuf<int32_t> flag;
    flag.alloc(N);
    m._node_detect.assign(N, -1);
    m._n_moved = 0;
    if (ni > 0) {
        hipLaunchKernelGGL(k_pp_nodes, dim3(pp_grid(N)), dim3(PP_THREADS), 0, 0, N, X[0].p, X[1].p, X[2].p, incl.p, ni,
                           m._opt.sphere_zlim_as_reference ? 1 : 0, flag.p);
        PPCHK(hipGetLastError());
        flag.down(m._node_detect.data(), N);
        bool moved = false;
        for (int64_t n = 0; n < N; ++n)
            if (m._node_detect[n] <= -2) {
                int times = 0;
                m._node_detect[n] = m.detect_node(n, times);
                moved = moved || times > 0;
            }
        if (moved) {
            for (int d = 0; d < 3; ++d) X[d].up(m._X[d].data(), N);
            flag.up(m._node_detect.data(), N);
        }
    }
    m._nodes_detected = true;
    double t1 = now_ms();
    phase_ms[0] = t1 - t0;

    // ---- elements ----
    m._cut.clear();
    m._cut_index.assign(NE, -1);
    if (ni == 0) {
        m._elements_detected = true;
        m._enrich_edge.clear();
        m._enriched = true;
        m.store_shape_functions();
        phase_ms[1] = phase_ms[2] = 0.0;
        phase_ms[3] = now_ms() - t1;
        return;
    }
    DBuf<int32_t> conn, emat, cutf, pos, cut_index;
    conn.up(m._conn.data(), 4 * NE);
    emat.up(m._elem_mat.data(), NE);
    cutf.alloc(NE);
    pos.alloc(NE + 1);
    cut_index.alloc(NE);
    DBuf<int> err;
    reset_err(err);
    hipLaunchKernelGGL(k_pp_elem_flags, dim3(pp_grid(NE)), dim3(PP_THREADS), 0, 0, NE, kind, conn.p, flag.p, emat.p,
                       incl.p, coh ? 1 : 0, cutf.p, err.p);
    PPCHK(hipGetLastError());
    if (read_err(err) != INT_MAX) {
        m.analyze_elements();  // raises the host path's error for the first such element
        throw Error(PGF_ERR_STATE, "device and host element detection disagree");
    }
    exclusive_scan<int32_t>(cutf.p, pos.p, NE);
    int32_t nc32 = 0;
    PPCHK(hipMemcpy(&nc32, pos.p + NE, sizeof(int32_t), hipMemcpyDeviceToHost));
    const int64_t NC = nc32;
    DBuf<CutElem> cut;
    cut.alloc(NC);
    hipLaunchKernelGGL(k_pp_cut_fill, dim3(pp_grid(NE)), dim3(PP_THREADS), 0, 0, NE, kind, conn.p, flag.p, emat.p,
                       incl.p, coh ? 1 : 0, m._opt.default_coh_law, cutf.p, pos.p, cut.p, cut_index.p);
    PPCHK(hipGetLastError());
    emat.down(m._elem_mat.data(), NE);
    cut_index.down(m._cut_index.data(), NE);
    m._elements_detected = true;
    double t2 = now_ms();
    phase_ms[1] = t2 - t1;

    // ---- enrichment nodes ----
    DBuf<uint64_t> keys, sorted, uniq;
    keys.alloc(4 * NC);
    sorted.alloc(4 * NC);
    uniq.alloc(4 * NC);
    DBuf<int> nsel;
    nsel.alloc(1);
    int64_t n_init = 0;
    if (NC > 0) {
        hipLaunchKernelGGL(k_pp_keys, dim3(pp_grid(NC)), dim3(PP_THREADS), 0, 0, NC, kind, conn.p, cut.p, keys.p);
        PPCHK(hipGetLastError());
        size_t bytes = 0;
        PPCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys.p, sorted.p, (int)(4 * NC)));
        DBuf<unsigned char> tmp;
        tmp.alloc(bytes);
        PPCHK(hipcub::DeviceRadixSort::SortKeys(tmp.p, bytes, keys.p, sorted.p, (int)(4 * NC)));
        size_t bytes2 = 0;
        PPCHK(hipcub::DeviceSelect::Unique(nullptr, bytes2, sorted.p, uniq.p, nsel.p, (int)(4 * NC)));
        DBuf<unsigned char> tmp2;
        tmp2.alloc(bytes2);
        PPCHK(hipcub::DeviceSelect::Unique(tmp2.p, bytes2, sorted.p, uniq.p, nsel.p, (int)(4 * NC)));
        int nu = 0;
        nsel.down(&nu, 1);
        uint64_t last = 0;
        if (nu > 0) PPCHK(hipMemcpy(&last, uniq.p + nu - 1, sizeof(uint64_t), hipMemcpyDeviceToHost));
        n_init = (nu > 0 && last == KEY_NONE) ? nu - 1 : nu;
    }
    const int64_t n_total = coh ? 2 * n_init : n_init;
    DBuf<double> XA[3];
    for (int d = 0; d < 3; ++d) {
        XA[d].alloc(N + n_total);
        PPCHK(hipMemcpy(XA[d].p, X[d].p, sizeof(double) * N, hipMemcpyDeviceToDevice));
    }
    if (n_init > 0) {
        reset_err(err);
        hipLaunchKernelGGL(k_pp_roots, dim3(pp_grid(n_init)), dim3(PP_THREADS), 0, 0, n_init, uniq.p, N, coh ? 1 : 0,
                           incl.p, XA[0].p, XA[1].p, XA[2].p, err.p);
        PPCHK(hipGetLastError());
        if (read_err(err) != INT_MAX) {
            m._cut.resize(NC);
            cut.down(m._cut.data(), NC);
            m.add_enrichments();  // raises the host path's error
            throw Error(PGF_ERR_STATE, "device and host enrichment disagree");
        }
        hipLaunchKernelGGL(k_pp_enrich_ids, dim3(pp_grid(NC)), dim3(PP_THREADS), 0, 0, NC, kind, conn.p, cut.p,
                           uniq.p, n_init, coh ? 1 : 0);
        PPCHK(hipGetLastError());
    }
    m._cut.resize(NC);
    cut.down(m._cut.data(), NC);
    std::vector<uint64_t> hk(n_init);
    if (n_init) uniq.down(hk.data(), n_init);
    m._enrich_edge.resize(n_total);
    for (int64_t k = 0; k < n_init; ++k) {
        std::array<int32_t, 3> e = {(int32_t)(hk[k] >> (KEY_NODE_BITS + KEY_INCL_BITS)),
                                    (int32_t)((hk[k] >> KEY_INCL_BITS) & ((1ull << KEY_NODE_BITS) - 1)),
                                    (int32_t)(hk[k] & ((1ull << KEY_INCL_BITS) - 1))};
        m._enrich_edge[k] = e;
        if (coh) m._enrich_edge[n_init + k] = e;
    }
    for (int d = 0; d < 3; ++d) {
        m._X[d].resize(N + n_total);
        if (n_total) PPCHK(hipMemcpy(m._X[d].data() + N, XA[d].p + N, sizeof(double) * n_total, hipMemcpyDeviceToHost));
    }
    m._enriched = true;
    double t3 = now_ms();
    phase_ms[2] = t3 - t2;

    // ---- integration tables ----
    DBuf<int64_t> vc, cc, gc, vptr, cptr, gptr;
    vc.alloc(NC); cc.alloc(NC); gc.alloc(NC);
    vptr.alloc(NC + 1); cptr.alloc(NC + 1); gptr.alloc(NC + 1);
    if (NC > 0) {
        hipLaunchKernelGGL(k_pp_tab_sizes, dim3(pp_grid(NC)), dim3(PP_THREADS), 0, 0, NC, dim, cut.p, vc.p, cc.p, gc.p);
        PPCHK(hipGetLastError());
    }
    exclusive_scan<int64_t>(vc.p, vptr.p, NC);
    exclusive_scan<int64_t>(cc.p, cptr.p, NC);
    exclusive_scan<int64_t>(gc.p, gptr.p, NC);
    m.vpt_ptr.assign(NC + 1, 0);
    m.cpt_ptr.assign(NC + 1, 0);
    vptr.down(m.vpt_ptr.data(), NC + 1);
    cptr.down(m.cpt_ptr.data(), NC + 1);
    int64_t ng = 0;
    PPCHK(hipMemcpy(&ng, gptr.p + NC, sizeof(int64_t), hipMemcpyDeviceToHost));
    const int64_t NV = m.vpt_ptr[NC], NCP = m.cpt_ptr[NC];
    DBuf<double> vw, vgrad, cw, cN, cR;
    DBuf<int32_t> vmat, claw;
    DBuf<int64_t> vgoff;
    DBuf<uint8_t> cnodes, cnh;
    vw.alloc(NV); vmat.alloc(NV); vgoff.alloc(NV); vgrad.alloc(ng);
    cw.alloc(NCP); cN.alloc(4 * NCP); cR.alloc(9 * NCP); cnodes.alloc(8 * NCP); cnh.alloc(NCP); claw.alloc(NCP);
    DBuf<geom::PolygonTable> poly;
    poly.up(&geom::polygon_table(), 1);
    if (NC > 0) {
        reset_err(err);
        TabArgs a;
        a.nc = NC; a.n_nodes = N; a.dim = dim; a.kind = kind; a.tet_weight_scale = m._opt.tet_weight_scale;
        a.cut = cut.p; a.conn = conn.p; a.X0 = XA[0].p; a.X1 = XA[1].p; a.X2 = XA[2].p; a.poly = poly.p;
        a.vptr = vptr.p; a.cptr = cptr.p; a.gptr = gptr.p;
        a.vw = vw.p; a.vmat = vmat.p; a.vgoff = vgoff.p; a.vgrad = vgrad.p;
        a.cw = cw.p; a.cN = cN.p; a.cR = cR.p; a.cnodes = cnodes.p; a.cnh = cnh.p; a.claw = claw.p;
        a.err = err.p;
        hipLaunchKernelGGL(k_pp_tables, dim3(pp_grid(NC)), dim3(PP_THREADS), 0, 0, a);
        PPCHK(hipGetLastError());
        if (read_err(err) != INT_MAX) {
            m.store_shape_functions();  // raises the host path's error
            throw Error(PGF_ERR_STATE, "device and host integration tables disagree");
        }
    }
    m.vpt_w.resize(NV); m.vpt_mat.resize(NV); m.vpt_grad_off.resize(NV); m.vpt_grad.resize(ng);
    m.cpt_w.resize(NCP); m.cpt_N.resize(4 * NCP); m.cpt_R.resize(9 * NCP); m.cpt_nodes.resize(8 * NCP);
    m.cpt_nh.resize(NCP); m.cpt_law.resize(NCP);
    vw.down(m.vpt_w.data(), NV); vmat.down(m.vpt_mat.data(), NV); vgoff.down(m.vpt_grad_off.data(), NV);
    vgrad.down(m.vpt_grad.data(), ng);
    cw.down(m.cpt_w.data(), NCP); cN.down(m.cpt_N.data(), 4 * NCP); cR.down(m.cpt_R.data(), 9 * NCP);
    cnodes.down(m.cpt_nodes.data(), 8 * NCP); cnh.down(m.cpt_nh.data(), NCP); claw.down(m.cpt_law.data(), NCP);
    phase_ms[3] = now_ms() - t3;
}

void device_preprocess(Mesh& m, double* phase_ms) { DevicePreprocess::run(m, phase_ms); }

}  // namespace pgf
