// Learned initial guess for the Newton/PCG solves: a small 3D U-Net on a
// 64^3 voxel grid (inclusion indicator, damage, step residual, previous load
// increment -> displacement correction), trained and evaluated on the device
// with this file's own kernels.
//
//  * every convolution is an implicit GEMM on bf16 MFMA tiles
//    (v_mfma_f32_16x16x16bf16_1k, FP32 accumulation), one wavefront per
//    16-position x 16*NT-channel output tile;
//  * activations are stored channels-last in bf16; gradients are FP32 with a
//    bf16 copy as MFMA operand; weights are FP32 masters with bf16 copies per
//    step (forward layout [Cout][T*Cin], data-gradient layout [Cin][T*Cout]);
//  * Adam in FP32.
//
// Layers (64^3 x 8 input):
//   a1 = relu(conv3(x))            16 @ 64^3
//   a2 = relu(down2(a1))           32 @ 32^3
//   a3 = relu(conv3(a2))           32 @ 32^3
//   a4 = relu(down2(a3))           64 @ 16^3
//   a5 = relu(conv3(a4))           64 @ 16^3
//   a6 = relu(up2(a5) + a3)        32 @ 32^3
//   a7 = relu(up2(a6) + a1)        16 @ 64^3
//   y  = conv3(a7)                 16 @ 64^3 (channels 0..2 used)
// conv3: 3x3x3, stride 1, zero padding 1; down2: 2x2x2 stride 2; up2: its
// transpose (each output voxel has exactly one tap).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/pgfem_learn.h"

namespace pgfn {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

#ifndef PGF_CNN_GRID
#define PGF_CNN_GRID 64
#endif
constexpr int GRID = PGF_CNN_GRID;
constexpr int NVOX = GRID * GRID * GRID;
constexpr int IN_CH = 8;
constexpr int OUT_CH = 3;
enum Mode { CONV3 = 0, DOWN = 1, UP = 2 };
constexpr int N_LAYERS = 8;

struct LayerSpec {
    int mode, cin, cout, din;  // din: input grid size
};
constexpr LayerSpec SPECS[N_LAYERS] = {
    {CONV3, 8, 16, GRID},      {DOWN, 16, 32, GRID},     {CONV3, 32, 32, GRID / 2}, {DOWN, 32, 64, GRID / 2},
    {CONV3, 64, 64, GRID / 4}, {UP, 64, 32, GRID / 4},   {UP, 32, 16, GRID / 2},    {CONV3, 16, 16, GRID},
};
__host__ __device__ constexpr int taps(int mode) { return mode == CONV3 ? 27 : 8; }
__host__ __device__ constexpr int dout_of(int mode, int din) { return mode == DOWN ? din / 2 : (mode == UP ? din * 2 : din); }

__device__ __host__ inline bf16 f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (bf16)(u >> 16);
}
__device__ __host__ inline float bf2f(bf16 h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// ---------------------------------------------------------------------------
// Implicit-GEMM convolution, forward form.
//   out[p][n] = bias[n] + sum_{t,c} in[src(p,t)][c] * B[n][t*C + c]
// CONV3: src = p + off(t) - 1 (zero outside); DOWN: src = 2p + off(t);
// UP: tiles are grouped by output parity class cls, the only tap is t = cls
// and src = p / 2.  One wavefront computes 16 output positions (consecutive
// along x; for UP consecutive input positions of one parity class) times
// 16*NT output channels.  Epilogue: + bias, + skip (bf16, same shape as out),
// ReLU, store bf16 and/or FP32 (or add into FP32).
struct ConvArgs {
    int mode, C, N, din, dout;
    const bf16* in;      // [din^3][C]
    const bf16* B;       // [N][T*C]
    const float* bias;   // [N] or null
    const bf16* skip;    // [dout^3][N] or null
    int relu;
    bf16* out_bf;        // [dout^3][N] or null
    float* out_f;        // [dout^3][N] or null
    int accumulate;      // out_f += instead of =
};

__device__ inline int3 tap_off(int mode, int t) {
    if (mode == CONV3) return make_int3(t % 3, (t / 3) % 3, t / 9);
    return make_int3(t & 1, (t >> 1) & 1, t >> 2);
}

template <int NT>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int T = taps(a.mode);
    const int n_tiles_pos = a.mode == UP ? 8 * (a.din * a.din * a.din / 16) : a.dout * a.dout * a.dout / 16;
    const int n_tiles_ch = a.N / (16 * NT);
    if (wave >= n_tiles_pos * n_tiles_ch) return;
    const int tpos = wave / n_tiles_ch, tch = wave % n_tiles_ch;
    const int row = lane & 15, kq = lane >> 4;
    // output position of this lane's A row (m = row)
    int cls = 0, px, py, pz, ix = 0, iy = 0, iz = 0;
    if (a.mode == UP) {
        const int per = a.din * a.din * a.din / 16;
        cls = tpos / per;
        const int i = (tpos % per) * 16 + row;
        ix = i % a.din; iy = (i / a.din) % a.din; iz = i / (a.din * a.din);
        px = 2 * ix + (cls & 1); py = 2 * iy + ((cls >> 1) & 1); pz = 2 * iz + (cls >> 2);
    } else {
        const int p = tpos * 16 + row;
        px = p % a.dout; py = (p / a.dout) % a.dout; pz = p / (a.dout * a.dout);
    }
    const int K = a.mode == UP ? a.C : T * a.C;
    const int ldb = T * a.C;
    const int boff = a.mode == UP ? cls * a.C : 0;
    v4f acc[NT];
    for (int j = 0; j < NT; ++j) acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        const int kk = k0 + kq * 4;  // this lane's 4 consecutive k
        v4s av = v4s{0, 0, 0, 0};
        if (kk < K) {
            int src = -1, c;
            if (a.mode == UP) {
                c = kk;
                src = (iz * a.din + iy) * a.din + ix;
            } else {
                const int t = kk / a.C;
                c = kk - t * a.C;
                const int3 o = tap_off(a.mode, t);
                int sx, sy, sz;
                if (a.mode == CONV3) { sx = px + o.x - 1; sy = py + o.y - 1; sz = pz + o.z - 1; }
                else { sx = 2 * px + o.x; sy = 2 * py + o.y; sz = 2 * pz + o.z; }
                if (sx >= 0 && sy >= 0 && sz >= 0 && sx < a.din && sy < a.din && sz < a.din) src = (sz * a.din + sy) * a.din + sx;
            }
            if (src >= 0) {
                const uint2 v = *reinterpret_cast<const uint2*>(a.in + (size_t)src * a.C + c);
                av = v4s{(short)(v.x & 0xffff), (short)(v.x >> 16), (short)(v.y & 0xffff), (short)(v.y >> 16)};
            }
        }
        for (int j = 0; j < NT; ++j) {
            const int n = (tch * NT + j) * 16 + row;
            v4s bv = v4s{0, 0, 0, 0};
            if (kk < K) {
                const uint2 v = *reinterpret_cast<const uint2*>(a.B + (size_t)n * ldb + boff + kk);
                bv = v4s{(short)(v.x & 0xffff), (short)(v.x >> 16), (short)(v.y & 0xffff), (short)(v.y >> 16)};
            }
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[j], 0, 0, 0);
        }
    }
    // epilogue: lane holds rows m = kq*4 + i, column n = row
    for (int j = 0; j < NT; ++j) {
        const int n = (tch * NT + j) * 16 + row;
        const float bn = a.bias ? a.bias[n] : 0.f;
        for (int i = 0; i < 4; ++i) {
            const int m = kq * 4 + i;
            size_t p;
            if (a.mode == UP) {
                const int per = a.din * a.din * a.din / 16;
                const int ii = (tpos % per) * 16 + m;
                const int jx = ii % a.din, jy = (ii / a.din) % a.din, jz = ii / (a.din * a.din);
                p = ((size_t)(2 * jz + (cls >> 2)) * a.dout + 2 * jy + ((cls >> 1) & 1)) * a.dout + 2 * jx + (cls & 1);
            } else {
                p = (size_t)tpos * 16 + m;
            }
            float v = acc[j][i] + bn;
            if (a.skip) v += bf2f(a.skip[p * a.N + n]);
            if (a.relu) v = v > 0.f ? v : 0.f;
            if (a.out_bf) a.out_bf[p * a.N + n] = f2bf(v);
            if (a.out_f) {
                if (a.accumulate) a.out_f[p * a.N + n] += v;
                else a.out_f[p * a.N + n] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[t][c][n] = sum_p in[src(p,t)][c] * g[p][n] over the
// layer's output positions p, as MFMA tiles (M = n, N = t*C + c, K = p) over
// chunks of positions; partial tiles go to part[chunk][N_out][T*C], summed
// by k_wgrad_reduce in chunk order.
struct WgArgs {
    int mode, C, N, din, dout, chunk;
    const bf16* in;  // [din^3][C]
    const bf16* g;   // [dout^3][N]
    float* part;     // [n_chunks][N][T*C]
};

__global__ __launch_bounds__(256) void k_wgrad(WgArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int T = taps(a.mode);
    const int TC = T * a.C;
    const int ntn = a.N / 16, ntk = (TC + 15) / 16;
    const int np = a.dout * a.dout * a.dout;
    const int nchunks = np / a.chunk;
    if (wave >= nchunks * ntn * ntk) return;
    const int ch = wave / (ntn * ntk);
    const int rem = wave % (ntn * ntk);
    const int tn = rem / ntk, tk = rem % ntk;
    const int row = lane & 15, kq = lane >> 4;
    const int n = tn * 16 + row;              // A row (output channel)
    const int col = tk * 16 + row;            // B column (t*C + c)
    const int t = col / a.C, c = col - t * a.C;
    const int3 o = tap_off(a.mode, t < T ? t : 0);
    v4f acc = v4f{0.f, 0.f, 0.f, 0.f};
    for (int p0 = ch * a.chunk; p0 < (ch + 1) * a.chunk; p0 += 16) {
        short av[4], bv[4];
        for (int i = 0; i < 4; ++i) {
            const int p = p0 + kq * 4 + i;
            av[i] = (short)a.g[(size_t)p * a.N + n];
            short b = 0;
            if (col < TC) {
                const int x = p % a.dout, y = (p / a.dout) % a.dout, z = p / (a.dout * a.dout);
                int sx, sy, sz;
                bool ok;
                if (a.mode == CONV3) {
                    sx = x + o.x - 1; sy = y + o.y - 1; sz = z + o.z - 1;
                    ok = sx >= 0 && sy >= 0 && sz >= 0 && sx < a.din && sy < a.din && sz < a.din;
                } else if (a.mode == DOWN) {
                    sx = 2 * x + o.x; sy = 2 * y + o.y; sz = 2 * z + o.z;
                    ok = true;
                } else {
                    sx = x >> 1; sy = y >> 1; sz = z >> 1;
                    ok = (x & 1) == o.x && (y & 1) == o.y && (z & 1) == o.z;
                }
                if (ok) b = (short)a.in[((size_t)(sz * a.din + sy) * a.din + sx) * a.C + c];
            }
            bv[i] = b;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(v4s{av[0], av[1], av[2], av[3]}, v4s{bv[0], bv[1], bv[2], bv[3]},
                                                        acc, 0, 0, 0);
    }
    // lane holds rows (n) tn*16 + kq*4 + i, column tk*16 + row
    if (col < TC)
        for (int i = 0; i < 4; ++i) {
            const int nn = tn * 16 + kq * 4 + i;
            a.part[((size_t)ch * a.N + nn) * TC + col] = acc[i];
        }
}

// gW[t][c][n] = sum over chunks of part[chunk][n][t*C + c]
__global__ void k_wgrad_reduce(const float* part, int nchunks, int N, int TC, int C, float* gW) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * TC) return;
    const int n = idx / TC, col = idx % TC;
    float s = 0.f;
    for (int k = 0; k < nchunks; ++k) s += part[((size_t)k * N + n) * TC + col];
    const int t = col / C, c = col % C;
    gW[((size_t)t * C + c) * N + n] = s;
}

// bias gradient: gb[n] = sum_p g[p][n]; one block per channel, fixed order
__global__ void k_bias_grad(const float* g, int np, int N, float* gb) {
    __shared__ float sh[256];
    const int n = blockIdx.x;
    float s = 0.f;
    for (int p = threadIdx.x; p < np; p += blockDim.x) s += g[(size_t)p * N + n];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) gb[n] = sh[0];
}

// g = d * (a > 0) (ReLU mask from the stored activation); gbf = bf16(g)
__global__ void k_relu_back(int64_t n, const float* d, const bf16* act, float* g, bf16* gbf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = bf2f(act[i]) > 0.f ? d[i] : 0.f;
    g[i] = v;
    gbf[i] = f2bf(v);
}
__global__ void k_to_bf16(int64_t n, const float* x, bf16* y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// bf16 weight copies from the FP32 master W[t][c][n]:
//   Bf[n][t*C + c] = W[t][c][n]             (forward)
//   Bd[c][t*N + n] = W[flip(t)][c][n]       (data gradient; flip = 26 - t for conv3)
__global__ void k_weights_bf16(int mode, int T, int C, int N, const float* W, bf16* Bf, bf16* Bd) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= T * C * N) return;
    const int t = idx / (C * N), c = (idx / N) % C, n = idx % N;
    const bf16 h = f2bf(W[idx]);
    Bf[(size_t)n * T * C + t * C + c] = h;
    const int tf = mode == CONV3 ? 26 - t : t;
    Bd[(size_t)c * T * N + tf * N + n] = h;
}

// Adam (FP32): m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2; w -= lr mhat / (sqrt(vhat) + eps)
__global__ void k_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// MSE over the 3 used output channels: dy = 2 (y - t) / (3 NVOX); per-block loss partials
__global__ void k_mse(const float* y, const float* tgt, float* dy, float* part) {
    __shared__ float sh[256];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    float s = 0.f;
    if (p < NVOX)
        for (int c = 0; c < 16; ++c) {
            float d = 0.f;
            if (c < OUT_CH) {
                d = y[(size_t)p * 16 + c] - tgt[(size_t)p * OUT_CH + c];
                s += d * d;
            }
            dy[(size_t)p * 16 + c] = c < OUT_CH ? 2.f * d / (3.f * NVOX) : 0.f;
        }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

}  // namespace pgfn

// ===========================================================================
// Host side
// ===========================================================================
using namespace pgfn;

#define NCHK(x)                                                                      \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) throw std::string("HIP error: ") + hipGetErrorString(e_); \
    } while (0)

struct pgf_net {
    std::string err;
    hipStream_t stream = nullptr;
    int64_t n_params = 0;
    int64_t w_off[N_LAYERS], b_off[N_LAYERS];
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    bf16 *Bf[N_LAYERS] = {}, *Bd[N_LAYERS] = {};
    // activations (bf16), layer outputs a1..a7, input x, final y (fp32)
    bf16* act[N_LAYERS] = {};  // act[l] = output of layer l (l < 7); act input separately
    bf16* x = nullptr;
    float* y = nullptr;
    float *dbuf[N_LAYERS] = {}, *gbuf[N_LAYERS] = {};  // d (fp32 grad wrt layer output), g (after ReLU mask)
    bf16* gbf[N_LAYERS] = {};
    float* part = nullptr;
    size_t part_size = 0;
    float* lossp = nullptr;
    float* dy = nullptr;
    // dataset on the device
    std::vector<bf16*> ds_x;
    std::vector<float*> ds_t;
    int64_t step = 0;
    ~pgf_net();
};

static int64_t act_count(int l) {
    const int d = dout_of(SPECS[l].mode, SPECS[l].din);
    return (int64_t)d * d * d * SPECS[l].cout;
}

template <typename T>
static T* nalloc(size_t n) {
    void* p = nullptr;
    NCHK(hipMalloc(&p, sizeof(T) * (n ? n : 1)));
    return (T*)p;
}

pgf_net::~pgf_net() {
    void* ptrs[] = {params, grads, adam_m, adam_v, x, y, part, lossp, dy};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (int l = 0; l < N_LAYERS; ++l) {
        void* q[] = {Bf[l], Bd[l], act[l], dbuf[l], gbuf[l], gbf[l]};
        for (void* p : q)
            if (p) (void)hipFree(p);
    }
    for (auto p : ds_x) (void)hipFree(p);
    for (auto p : ds_t) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
}

static std::string g_net_err;

static void launch_conv(pgf_net* net, int mode, int C, int N, int din, const bf16* in, const bf16* B, const float* bias,
                        const bf16* skip, int relu, bf16* out_bf, float* out_f, int accumulate) {
    ConvArgs a{mode, C, N, din, dout_of(mode, din), in, B, bias, skip, relu, out_bf, out_f, accumulate};
    const int dout = a.dout;
    const int tiles_pos = mode == UP ? 8 * (din * din * din / 16) : dout * dout * dout / 16;
    const int NT = N >= 64 ? 4 : (N >= 32 ? 2 : 1);
    const int waves = tiles_pos * (N / (16 * NT));
    const int blocks = (waves + 3) / 4;
    if (NT == 4) hipLaunchKernelGGL(k_conv<4>, dim3(blocks), dim3(256), 0, net->stream, a);
    else if (NT == 2) hipLaunchKernelGGL(k_conv<2>, dim3(blocks), dim3(256), 0, net->stream, a);
    else hipLaunchKernelGGL(k_conv<1>, dim3(blocks), dim3(256), 0, net->stream, a);
    NCHK(hipGetLastError());
}

static void launch_wgrad(pgf_net* net, int l, const bf16* in, const bf16* g, float* gW) {
    const LayerSpec& s = SPECS[l];
    const int T = taps(s.mode), TC = T * s.cin;
    const int dout = dout_of(s.mode, s.din);
    const int np = dout * dout * dout;
    const int chunk = std::min(np >= 65536 ? 1024 : 256, np);
    const int nchunks = np / chunk;
    WgArgs a{s.mode, s.cin, s.cout, s.din, dout, chunk, in, g, net->part};
    const int waves = nchunks * (s.cout / 16) * ((TC + 15) / 16);
    hipLaunchKernelGGL(k_wgrad, dim3((waves + 3) / 4), dim3(256), 0, net->stream, a);
    NCHK(hipGetLastError());
    const int n = s.cout * TC;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((n + 255) / 256), dim3(256), 0, net->stream, (const float*)net->part, nchunks,
                       s.cout, TC, s.cin, gW);
    NCHK(hipGetLastError());
}

static void refresh_bf16(pgf_net* net) {
    for (int l = 0; l < N_LAYERS; ++l) {
        const LayerSpec& s = SPECS[l];
        const int n = taps(s.mode) * s.cin * s.cout;
        hipLaunchKernelGGL(k_weights_bf16, dim3((n + 255) / 256), dim3(256), 0, net->stream, s.mode, taps(s.mode), s.cin,
                           s.cout, (const float*)(net->params + net->w_off[l]), net->Bf[l], net->Bd[l]);
    }
    NCHK(hipGetLastError());
}

static void forward(pgf_net* net, const bf16* x) {
    float* P = net->params;
    // a1 .. a7
    launch_conv(net, CONV3, 8, 16, GRID, x, net->Bf[0], P + net->b_off[0], nullptr, 1, net->act[0], nullptr, 0);
    launch_conv(net, DOWN, 16, 32, GRID, net->act[0], net->Bf[1], P + net->b_off[1], nullptr, 1, net->act[1], nullptr, 0);
    launch_conv(net, CONV3, 32, 32, GRID / 2, net->act[1], net->Bf[2], P + net->b_off[2], nullptr, 1, net->act[2], nullptr, 0);
    launch_conv(net, DOWN, 32, 64, GRID / 2, net->act[2], net->Bf[3], P + net->b_off[3], nullptr, 1, net->act[3], nullptr, 0);
    launch_conv(net, CONV3, 64, 64, GRID / 4, net->act[3], net->Bf[4], P + net->b_off[4], nullptr, 1, net->act[4], nullptr, 0);
    launch_conv(net, UP, 64, 32, GRID / 4, net->act[4], net->Bf[5], P + net->b_off[5], net->act[2], 1, net->act[5], nullptr, 0);
    launch_conv(net, UP, 32, 16, GRID / 2, net->act[5], net->Bf[6], P + net->b_off[6], net->act[0], 1, net->act[6], nullptr, 0);
    launch_conv(net, CONV3, 16, 16, GRID, net->act[6], net->Bf[7], P + net->b_off[7], nullptr, 0, nullptr, net->y, 0);
}

static void relu_back(pgf_net* net, int l) {
    const int64_t n = act_count(l);
    hipLaunchKernelGGL(k_relu_back, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, net->stream, n,
                       (const float*)net->dbuf[l], (const bf16*)net->act[l], net->gbuf[l], net->gbf[l]);
    NCHK(hipGetLastError());
}

static void bias_grad(pgf_net* net, int l, const float* g) {
    const LayerSpec& s = SPECS[l];
    const int d = dout_of(s.mode, s.din);
    hipLaunchKernelGGL(k_bias_grad, dim3(s.cout), dim3(256), 0, net->stream, g, d * d * d, s.cout,
                       net->grads + net->b_off[l]);
    NCHK(hipGetLastError());
}

// gradients of sum(dy . y) with respect to every parameter, dy [NVOX][16] FP32 on the device
static void backward(pgf_net* net, const bf16* x, const float* dy) {
    float* G = net->grads;
    const int64_t ny = (int64_t)NVOX * 16;
    // layer 7 (output conv3, no ReLU): g7 = dy
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((ny + 255) / 256)), dim3(256), 0, net->stream, ny, dy, net->gbf[7]);
    launch_wgrad(net, 7, net->act[6], net->gbf[7], G + net->w_off[7]);
    bias_grad(net, 7, dy);
    // d a7 = conv3^T(g7)
    launch_conv(net, CONV3, 16, 16, GRID, net->gbf[7], net->Bd[7], nullptr, nullptr, 0, nullptr, net->dbuf[6], 0);
    relu_back(net, 6);  // g6 (layer 6 = up2 into a7, skip from a1)
    launch_wgrad(net, 6, net->act[5], net->gbf[6], G + net->w_off[6]);
    bias_grad(net, 6, net->gbuf[6]);
    // d a6 = up2^T(g6) (down form); skip: d a1 gets g6 (fp32)
    launch_conv(net, DOWN, 16, 32, GRID, net->gbf[6], net->Bd[6], nullptr, nullptr, 0, nullptr, net->dbuf[5], 0);
    relu_back(net, 5);
    launch_wgrad(net, 5, net->act[4], net->gbf[5], G + net->w_off[5]);
    bias_grad(net, 5, net->gbuf[5]);
    launch_conv(net, DOWN, 32, 64, GRID / 2, net->gbf[5], net->Bd[5], nullptr, nullptr, 0, nullptr, net->dbuf[4], 0);
    relu_back(net, 4);
    launch_wgrad(net, 4, net->act[3], net->gbf[4], G + net->w_off[4]);
    bias_grad(net, 4, net->gbuf[4]);
    launch_conv(net, CONV3, 64, 64, GRID / 4, net->gbf[4], net->Bd[4], nullptr, nullptr, 0, nullptr, net->dbuf[3], 0);
    relu_back(net, 3);
    launch_wgrad(net, 3, net->act[2], net->gbf[3], G + net->w_off[3]);
    bias_grad(net, 3, net->gbuf[3]);
    // d a3 = down2^T(g3) (up form) + g5 (skip into a6)
    NCHK(hipMemcpyAsync(net->dbuf[2], net->gbuf[5], sizeof(float) * act_count(2), hipMemcpyDeviceToDevice, net->stream));
    launch_conv(net, UP, 64, 32, GRID / 4, net->gbf[3], net->Bd[3], nullptr, nullptr, 0, nullptr, net->dbuf[2], 1);
    relu_back(net, 2);
    launch_wgrad(net, 2, net->act[1], net->gbf[2], G + net->w_off[2]);
    bias_grad(net, 2, net->gbuf[2]);
    launch_conv(net, CONV3, 32, 32, GRID / 2, net->gbf[2], net->Bd[2], nullptr, nullptr, 0, nullptr, net->dbuf[1], 0);
    relu_back(net, 1);
    launch_wgrad(net, 1, net->act[0], net->gbf[1], G + net->w_off[1]);
    bias_grad(net, 1, net->gbuf[1]);
    // d a1 = down2^T(g1) (up form) + g6 (skip into a7)
    NCHK(hipMemcpyAsync(net->dbuf[0], net->gbuf[6], sizeof(float) * act_count(0), hipMemcpyDeviceToDevice, net->stream));
    launch_conv(net, UP, 32, 16, GRID / 2, net->gbf[1], net->Bd[1], nullptr, nullptr, 0, nullptr, net->dbuf[0], 1);
    relu_back(net, 0);
    launch_wgrad(net, 0, x, net->gbf[0], G + net->w_off[0]);
    bias_grad(net, 0, net->gbuf[0]);
}

template <typename F>
static int nguard(pgf_net* net, F&& f) {
    try {
        f();
        return 0;
    } catch (const std::string& s) {
        if (net) net->err = s;
        else g_net_err = s;
        return PGF_EDEVICE;
    }
}

static void upload_x(pgf_net* net, const float* xh, bf16* dst) {
    std::vector<bf16> h((size_t)NVOX * IN_CH);
    for (size_t i = 0; i < h.size(); ++i) h[i] = f2bf(xh[i]);
    NCHK(hipMemcpyAsync(dst, h.data(), sizeof(bf16) * h.size(), hipMemcpyHostToDevice, net->stream));
    NCHK(hipStreamSynchronize(net->stream));
}

extern "C" {

PGF_API int pgf_net_create(uint64_t seed, pgf_net** out) {
    if (!out) return PGF_EARG;
    pgf_net* net = new pgf_net();
    int rc = nguard(nullptr, [&]() {
        NCHK(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));
        int64_t off = 0;
        for (int l = 0; l < N_LAYERS; ++l) {
            const LayerSpec& s = SPECS[l];
            net->w_off[l] = off;
            off += (int64_t)taps(s.mode) * s.cin * s.cout;
            net->b_off[l] = off;
            off += s.cout;
        }
        net->n_params = off;
        // He-normal initialisation from a fixed 64-bit LCG + Box-Muller; biases 0
        std::vector<float> h(off, 0.f);
        uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
        auto uni = [&]() {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            return ((st >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        };
        for (int l = 0; l < N_LAYERS; ++l) {
            const LayerSpec& s = SPECS[l];
            const int fan_in = (s.mode == UP ? 1 : taps(s.mode)) * s.cin;
            const double sd = std::sqrt(2.0 / fan_in) * (l == N_LAYERS - 1 ? 0.1 : 1.0);
            for (int64_t i = 0; i < (int64_t)taps(s.mode) * s.cin * s.cout; ++i) {
                const double u1 = uni(), u2 = uni();
                h[net->w_off[l] + i] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
            }
        }
        net->params = nalloc<float>(off);
        net->grads = nalloc<float>(off);
        net->adam_m = nalloc<float>(off);
        net->adam_v = nalloc<float>(off);
        NCHK(hipMemcpy(net->params, h.data(), sizeof(float) * off, hipMemcpyHostToDevice));
        NCHK(hipMemset(net->adam_m, 0, sizeof(float) * off));
        NCHK(hipMemset(net->adam_v, 0, sizeof(float) * off));
        NCHK(hipMemset(net->grads, 0, sizeof(float) * off));
        size_t part = 0;
        for (int l = 0; l < N_LAYERS; ++l) {
            const LayerSpec& s = SPECS[l];
            const int T = taps(s.mode);
            net->Bf[l] = nalloc<bf16>((size_t)T * s.cin * s.cout);
            net->Bd[l] = nalloc<bf16>((size_t)T * s.cin * s.cout);
            net->act[l] = nalloc<bf16>(act_count(l));
            net->dbuf[l] = nalloc<float>(act_count(l));
            net->gbuf[l] = nalloc<float>(act_count(l));
            net->gbf[l] = nalloc<bf16>(act_count(l));
            const int d = dout_of(s.mode, s.din);
            const int np = d * d * d;
            const int chunk = std::min(np >= 65536 ? 1024 : 256, np);
            part = std::max(part, (size_t)(np / chunk) * s.cout * (size_t)T * s.cin);
        }
        net->part = nalloc<float>(part);
        net->part_size = part;
        net->x = nalloc<bf16>((size_t)NVOX * IN_CH);
        net->y = nalloc<float>((size_t)NVOX * 16);
        net->dy = nalloc<float>((size_t)NVOX * 16);
        net->lossp = nalloc<float>(NVOX / 256);
        refresh_bf16(net);
        NCHK(hipStreamSynchronize(net->stream));
    });
    if (rc) {
        delete net;
        return rc;
    }
    *out = net;
    return 0;
}

PGF_API void pgf_net_destroy(pgf_net* net) { delete net; }

PGF_API const char* pgf_net_last_error(const pgf_net* net) { return net ? net->err.c_str() : g_net_err.c_str(); }

PGF_API int pgf_net_num_params(pgf_net* net, int64_t* n) {
    if (!net || !n) return PGF_EARG;
    *n = net->n_params;
    return 0;
}

PGF_API int pgf_net_get_params(pgf_net* net, float* w) {
    if (!net || !w) return PGF_EARG;
    return nguard(net, [&]() { NCHK(hipMemcpy(w, net->params, sizeof(float) * net->n_params, hipMemcpyDeviceToHost)); });
}

PGF_API int pgf_net_set_params(pgf_net* net, const float* w) {
    if (!net || !w) return PGF_EARG;
    return nguard(net, [&]() {
        NCHK(hipMemcpy(net->params, w, sizeof(float) * net->n_params, hipMemcpyHostToDevice));
        refresh_bf16(net);
        NCHK(hipStreamSynchronize(net->stream));
    });
}

PGF_API int pgf_net_forward(pgf_net* net, const float* x, float* y) {
    if (!net || !x || !y) return PGF_EARG;
    return nguard(net, [&]() {
        upload_x(net, x, net->x);
        forward(net, net->x);
        std::vector<float> h((size_t)NVOX * 16);
        NCHK(hipMemcpyAsync(h.data(), net->y, sizeof(float) * h.size(), hipMemcpyDeviceToHost, net->stream));
        NCHK(hipStreamSynchronize(net->stream));
        for (int64_t p = 0; p < NVOX; ++p)
            for (int c = 0; c < OUT_CH; ++c) y[p * OUT_CH + c] = h[p * 16 + c];
    });
}

PGF_API int pgf_net_backward(pgf_net* net, const float* x, const float* dy, float* grads) {
    if (!net || !x || !dy || !grads) return PGF_EARG;
    return nguard(net, [&]() {
        upload_x(net, x, net->x);
        forward(net, net->x);
        std::vector<float> h((size_t)NVOX * 16, 0.f);
        for (int64_t p = 0; p < NVOX; ++p)
            for (int c = 0; c < OUT_CH; ++c) h[p * 16 + c] = dy[p * OUT_CH + c];
        NCHK(hipMemcpyAsync(net->dy, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, net->stream));
        backward(net, net->x, net->dy);
        NCHK(hipMemcpyAsync(grads, net->grads, sizeof(float) * net->n_params, hipMemcpyDeviceToHost, net->stream));
        NCHK(hipStreamSynchronize(net->stream));
    });
}

PGF_API int pgf_net_add_sample(pgf_net* net, const float* x, const float* t) {
    if (!net || !x || !t) return PGF_EARG;
    return nguard(net, [&]() {
        bf16* dx = nalloc<bf16>((size_t)NVOX * IN_CH);
        float* dt = nalloc<float>((size_t)NVOX * OUT_CH);
        net->ds_x.push_back(dx);
        net->ds_t.push_back(dt);
        upload_x(net, x, dx);
        NCHK(hipMemcpy(dt, t, sizeof(float) * NVOX * OUT_CH, hipMemcpyHostToDevice));
    });
}

PGF_API int pgf_net_num_samples(pgf_net* net, int32_t* n) {
    if (!net || !n) return PGF_EARG;
    *n = (int32_t)net->ds_x.size();
    return 0;
}

// `epochs` passes over the device dataset in sample order, one Adam step per
// sample; loss[e] = mean MSE of epoch e (before each step's update)
PGF_API int pgf_net_train(pgf_net* net, int32_t epochs, float lr, float* loss) {
    if (!net || epochs < 0) return PGF_EARG;
    if (net->ds_x.empty()) return PGF_ESTATE;
    return nguard(net, [&]() {
        const int nb = NVOX / 256;
        std::vector<float> lp(nb);
        for (int e = 0; e < epochs; ++e) {
            double tot = 0.0;
            for (size_t s = 0; s < net->ds_x.size(); ++s) {
                forward(net, net->ds_x[s]);
                hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, net->stream, (const float*)net->y,
                                   (const float*)net->ds_t[s], net->dy, net->lossp);
                backward(net, net->ds_x[s], net->dy);
                ++net->step;
                const float b1 = 0.9f, b2 = 0.999f;
                const float c1 = 1.f - std::pow(b1, (float)net->step), c2 = 1.f - std::pow(b2, (float)net->step);
                hipLaunchKernelGGL(k_adam, dim3((unsigned)((net->n_params + 255) / 256)), dim3(256), 0, net->stream,
                                   net->n_params, net->params, (const float*)net->grads, net->adam_m, net->adam_v, lr, b1, b2,
                                   1e-8f, c1, c2);
                refresh_bf16(net);
                NCHK(hipMemcpyAsync(lp.data(), net->lossp, sizeof(float) * nb, hipMemcpyDeviceToHost, net->stream));
                NCHK(hipStreamSynchronize(net->stream));
                double s2 = 0.0;
                for (float v : lp) s2 += v;
                tot += s2 / (3.0 * NVOX);
            }
            if (loss) loss[e] = (float)(tot / net->ds_x.size());
        }
    });
}

}  // extern "C"

// ===========================================================================
// Learned initial guess inside the Newton step
// ===========================================================================
#include "pgf_device.hpp"

namespace pgfn {

// voxel of a point of the box [lo, lo + ext)
__device__ inline int vox_of(const double* X, int64_t n_all, int64_t node, const double* lo, const double* ext) {
    int v[3];
    for (int d = 0; d < 3; ++d) {
        int i = (int)floor((X[d * n_all + node] - lo[d]) / ext[d] * GRID);
        v[d] = i < 0 ? 0 : (i >= GRID ? GRID - 1 : i);
    }
    return (v[2] * GRID + v[1]) * GRID + v[0];
}

struct Box {
    double lo[3], ext[3];
};

__global__ void k_node_vox(int64_t n_all, const double* X, Box b, int32_t* node_vox) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < n_all) node_vox[n] = vox_of(X, n_all, n, b.lo, b.ext);
}

// cohesive point -> voxel of its cut element's centroid (regular slots 0..3)
__global__ void k_cpt_vox(int64_t n_cut, int dim, int64_t n_all, const double* X, const int32_t* ce_nodes, const int64_t* cpt_ptr,
                          Box b, int32_t* cpt_vox) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cut) return;
    const int nv = dim == 2 ? 4 : 4;
    int v[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) {
        double s = 0.0;
        for (int k = 0; k < nv; ++k) s += X[d * n_all + ce_nodes[c * 16 + k]];
        int i = (int)floor((s / nv - b.lo[d]) / b.ext[d] * GRID);
        v[d] = i < 0 ? 0 : (i >= GRID ? GRID - 1 : i);
    }
    const int vox = (v[2] * GRID + v[1]) * GRID + v[0];
    for (int64_t q = cpt_ptr[c]; q < cpt_ptr[c + 1]; ++q) cpt_vox[q] = vox;
}

__global__ void k_feat_init(const float* ind, float* feat, float* cnt) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= NVOX) return;
    feat[(size_t)v * IN_CH + 0] = ind ? ind[v] : 0.f;
    for (int c = 1; c < IN_CH; ++c) feat[(size_t)v * IN_CH + c] = 0.f;
    cnt[v] = 0.f;
}

// damage channel: max over the voxel's cohesive points of delta_max / delta_c
__global__ void k_feat_damage(int64_t n_cpt, const int32_t* cpt_vox, const double* dmax, const int32_t* cpt_law,
                              const pgfk::CohParams* laws, float* feat) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_cpt) return;
    const float val = (float)(dmax[q] / laws[cpt_law[q]].dc);
    // non-negative floats order like their bit patterns
    atomicMax(reinterpret_cast<int*>(&feat[(size_t)cpt_vox[q] * IN_CH + 1]), __float_as_int(val > 0.f ? val : 0.f));
}

// residual channels: sum of the free-dof residual of the nodes in each voxel
__global__ void k_feat_res(int64_t nf, int64_t n_all, const int64_t* free_slot, const double* r, const int32_t* node_vox,
                           float* feat) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int64_t s = free_slot[f];
    const int d = (int)(s / n_all);
    atomicAdd(&feat[(size_t)node_vox[s - d * n_all] * IN_CH + 2 + d], (float)r[f]);
}

// averaged nodal field a - b (or a alone when b is null) into channels ch..ch+dim-1
__global__ void k_feat_nodal(int64_t n_all, int dim, const double* a, const double* b, const int32_t* node_vox, float* out,
                             int ld, int ch, float* cnt) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_all) return;
    const int v = node_vox[n];
    for (int d = 0; d < dim; ++d) {
        const double x = a[d * n_all + n] - (b ? b[d * n_all + n] : 0.0);
        atomicAdd(&out[(size_t)v * ld + ch + d], (float)x);
    }
    atomicAdd(&cnt[v], 1.f);
}

__global__ void k_feat_avg(float* out, int ld, int ch, int nc, const float* cnt) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= NVOX) return;
    const float c = cnt[v];
    for (int k = 0; k < nc; ++k) out[(size_t)v * ld + ch + k] = c > 0.f ? out[(size_t)v * ld + ch + k] / c : 0.f;
}

// sum of squares of channels ch..ch+nc-1 (one double per block)
__global__ void k_sumsq(const float* x, int ld, int ch, int nc, double* part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < NVOX; v += gridDim.x * blockDim.x)
        for (int k = 0; k < nc; ++k) {
            const double t = x[(size_t)v * ld + ch + k];
            s += t * t;
        }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_scale_ch(float* x, int ld, int ch, int nc, float s) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= NVOX) return;
    for (int k = 0; k < nc; ++k) x[(size_t)v * ld + ch + k] *= s;
}

// Z columns on the free dofs: j < 3 the network output, channel j, trilinear
// at the node (cell-centred grid) on dofs of component j, zero elsewhere;
// column 3 the previous load increment.
__global__ void k_prolong(int64_t nf, int64_t n_all, const int64_t* free_slot, const double* X, Box b, const float* y,
                          int ldy, int use_net, const double* up, const double* upp, double* Z) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int64_t s = free_slot[f];
    const int d = (int)(s / n_all);
    const int64_t node = s - d * n_all;
    double val = 0.0;
    if (use_net && d < 3) {
        int i0[3];
        double w[3];
        for (int k = 0; k < 3; ++k) {
            double t = (X[k * n_all + node] - b.lo[k]) / b.ext[k] * GRID - 0.5;
            int i = (int)floor(t);
            if (i < 0) { i = 0; t = 0.0; }
            if (i > GRID - 2) { i = GRID - 2; t = GRID - 1; }
            i0[k] = i;
            w[k] = t - i;
        }
        for (int c = 0; c < 8; ++c) {
            const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
            const double wc = (dx ? w[0] : 1.0 - w[0]) * (dy ? w[1] : 1.0 - w[1]) * (dz ? w[2] : 1.0 - w[2]);
            const int v = ((i0[2] + dz) * GRID + i0[1] + dy) * GRID + i0[0] + dx;
            val += wc * y[(size_t)v * ldy + d];
        }
    }
    for (int j = 0; j < 3; ++j) Z[j * nf + f] = (j == d) ? val : 0.0;
    Z[3 * nf + f] = up ? up[s] - upp[s] : 0.0;
}

// y = K_ff x for one of the four columns (K in FP64 or FP32, vectors FP64)
template <typename T>
__global__ void k_spmv_col(int64_t nf, const int32_t* rowptr, const int32_t* rowsplit, const int32_t* colidx, const T* K,
                           const double* x, double* y) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nf) return;
    double s = 0.0;
    for (int32_t j = rowptr[r]; j < rowptr[r] + rowsplit[r]; ++j) s += (double)K[j] * x[colidx[j]];
    y[r] = s;
}

// 14 dot products: G_ij = Z_i . KZ_j (i <= j, 10) and b_i = Z_i . r (4); per-block partials
constexpr int NDOT = 14;
__global__ void k_dots(int64_t nf, const double* Z, const double* KZ, const double* r, double* part) {
    __shared__ double sh[NDOT][256];
    double acc[NDOT];
    for (int k = 0; k < NDOT; ++k) acc[k] = 0.0;
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += (int64_t)gridDim.x * blockDim.x) {
        double z[4], kz[4];
        for (int j = 0; j < 4; ++j) { z[j] = Z[j * nf + f]; kz[j] = KZ[j * nf + f]; }
        int k = 0;
        for (int i = 0; i < 4; ++i)
            for (int j = i; j < 4; ++j) acc[k++] += z[i] * kz[j];
        for (int i = 0; i < 4; ++i) acc[10 + i] += z[i] * r[f];
    }
    for (int k = 0; k < NDOT; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < NDOT; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < NDOT) part[threadIdx.x * gridDim.x + blockIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_combine(int64_t nf, const double* Z, double a0, double a1, double a2, double a3, double* x) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nf) x[f] = a0 * Z[f] + a1 * Z[nf + f] + a2 * Z[2 * nf + f] + a3 * Z[3 * nf + f];
}

}  // namespace pgfn

namespace pgfd {

struct LearnState {
    pgf_net* net = nullptr;
    pgfn::Box box{};
    float* ind = nullptr;      // [NVOX] indicator
    float* feat = nullptr;     // [NVOX][8] step-start features
    float* cnt = nullptr;      // [NVOX]
    float* tgt = nullptr;      // [NVOX][3]
    int32_t* node_vox = nullptr;
    int32_t* cpt_vox = nullptr;
    double *U_prev = nullptr, *U_prevprev = nullptr, *U_start = nullptr, *U_iter = nullptr;
    int n_commits = 0;
    bool have_features = false;
    double* Z = nullptr;       // [4][n_free]
    double* KZ = nullptr;
    double* part = nullptr;    // [NDOT][DOT_BLOCKS]
    double* h_part = nullptr;
    double last_alpha[4] = {0, 0, 0, 0};
    int last_cols = 0;
    double last_ms = 0.0;
};
constexpr int DOT_BLOCKS = 512;

void learn_free(LearnState* s) {
    if (!s) return;
    void* p[] = {s->ind, s->feat, s->cnt, s->tgt, s->node_vox, s->cpt_vox, s->U_prev, s->U_prevprev, s->U_start, s->U_iter,
                 s->Z, s->KZ, s->part};
    for (void* q : p)
        if (q) (void)hipFree(q);
    if (s->h_part) (void)hipHostFree(s->h_part);
    delete s;
}

static unsigned gfor(int64_t n) { return (unsigned)((n + 255) / 256); }

// normalise channels ch..ch+nc-1 of x [NVOX][ld] to unit RMS (left as is when zero)
static void normalise(Ctx& c, LearnState& L, float* x, int ld, int ch, int nc) {
    hipLaunchKernelGGL(pgfn::k_sumsq, dim3(DOT_BLOCKS), dim3(256), 0, c.stream, (const float*)x, ld, ch, nc, L.part);
    NCHK(hipMemcpyAsync(L.h_part, L.part, sizeof(double) * DOT_BLOCKS, hipMemcpyDeviceToHost, c.stream));
    NCHK(hipStreamSynchronize(c.stream));
    double s = 0.0;
    for (int b = 0; b < DOT_BLOCKS; ++b) s += L.h_part[b];
    const double rms = std::sqrt(s / ((double)pgfn::NVOX * nc));
    if (rms > 0.0)
        hipLaunchKernelGGL(pgfn::k_scale_ch, dim3(gfor(pgfn::NVOX)), dim3(256), 0, c.stream, x, ld, ch, nc, (float)(1.0 / rms));
}

// features at the start of a load step (or of a later Newton iteration: the
// residual channels then hold that iteration's residual): indicator, damage,
// residual, previous increment
static void build_features(Ctx& c, LearnState& L, bool step_start) {
    using namespace pgfn;
    hipLaunchKernelGGL(k_feat_init, dim3(gfor(NVOX)), dim3(256), 0, c.stream, (const float*)L.ind, L.feat, L.cnt);
    if (c.n_cpt > 0)
        hipLaunchKernelGGL(k_feat_damage, dim3(gfor(c.n_cpt)), dim3(256), 0, c.stream, c.n_cpt, (const int32_t*)L.cpt_vox,
                           (const double*)c.dmax_conv, (const int32_t*)c.cpt_law, (const pgfk::CohParams*)c.laws, L.feat);
    hipLaunchKernelGGL(k_feat_res, dim3(gfor(c.n_free)), dim3(256), 0, c.stream, c.n_free, c.n_all,
                       (const int64_t*)c.free_slot, (const double*)c.r, (const int32_t*)L.node_vox, L.feat);
    if (L.n_commits >= 2) {
        hipLaunchKernelGGL(k_feat_nodal, dim3(gfor(c.n_all)), dim3(256), 0, c.stream, c.n_all, c.dim, (const double*)L.U_prev,
                           (const double*)L.U_prevprev, (const int32_t*)L.node_vox, L.feat, IN_CH, 5, L.cnt);
        hipLaunchKernelGGL(k_feat_avg, dim3(gfor(NVOX)), dim3(256), 0, c.stream, L.feat, IN_CH, 5, c.dim, (const float*)L.cnt);
    }
    NCHK(hipGetLastError());
    normalise(c, L, L.feat, IN_CH, 2, 3);
    normalise(c, L, L.feat, IN_CH, 5, 3);
    if (step_start) {
        NCHK(hipMemcpyAsync(L.U_start, c.U, sizeof(double) * c.dim * c.n_all, hipMemcpyDeviceToDevice, c.stream));
        L.have_features = true;
    }
}

bool learn_guess(Ctx& c, int use_guess, bool fp32, int it) {
    using namespace pgfn;
    LearnState& L = *c.learn;
    hipEvent_t e0, e1;
    NCHK(hipEventCreate(&e0));
    NCHK(hipEventCreate(&e1));
    NCHK(hipEventRecord(e0, c.stream));
    const bool guess = use_guess == 2 || use_guess == 3;
    if (it == 0 || (guess && use_guess == 2)) build_features(c, L, it == 0);
    // column 3: the previous load increment at iteration 0, the previous Newton
    // correction U - U_iter afterwards
    const double* c3a = nullptr;
    const double* c3b = nullptr;
    if (it == 0) {
        if (L.n_commits >= 2) { c3a = L.U_prev; c3b = L.U_prevprev; }
    } else {
        c3a = c.U;
        c3b = L.U_iter;
    }
    if (!guess) {
        NCHK(hipMemcpyAsync(L.U_iter, c.U, sizeof(double) * c.dim * c.n_all, hipMemcpyDeviceToDevice, c.stream));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return false;
    }
    const bool use_net = use_guess == 2 && L.net != nullptr;
    if (use_net) {
        hipLaunchKernelGGL(k_to_bf16, dim3(gfor((int64_t)NVOX * IN_CH)), dim3(256), 0, c.stream, (int64_t)NVOX * IN_CH,
                           (const float*)L.feat, L.net->x);
        // the network runs on the solver's stream
        hipStream_t keep = L.net->stream;
        L.net->stream = c.stream;
        forward(L.net, L.net->x);
        L.net->stream = keep;
    }
    const bool have_prev = c3a != nullptr;
    const int64_t nf = c.n_free;
    hipLaunchKernelGGL(k_prolong, dim3(gfor(nf)), dim3(256), 0, c.stream, nf, c.n_all, (const int64_t*)c.free_slot,
                       (const double*)c.X, L.box, (const float*)(use_net ? L.net->y : nullptr), 16, use_net ? 1 : 0, c3a, c3b,
                       L.Z);
    NCHK(hipMemcpyAsync(L.U_iter, c.U, sizeof(double) * c.dim * c.n_all, hipMemcpyDeviceToDevice, c.stream));
    for (int j = 0; j < 4; ++j) {
        if (fp32)
            hipLaunchKernelGGL(k_spmv_col<float>, dim3(gfor(nf)), dim3(256), 0, c.stream, nf, (const int32_t*)c.rowptr,
                               (const int32_t*)c.rowsplit, (const int32_t*)c.colidx, (const float*)c.K32,
                               (const double*)(L.Z + j * nf), L.KZ + j * nf);
        else
            hipLaunchKernelGGL(k_spmv_col<double>, dim3(gfor(nf)), dim3(256), 0, c.stream, nf, (const int32_t*)c.rowptr,
                               (const int32_t*)c.rowsplit, (const int32_t*)c.colidx, (const double*)c.K,
                               (const double*)(L.Z + j * nf), L.KZ + j * nf);
    }
    hipLaunchKernelGGL(k_dots, dim3(DOT_BLOCKS), dim3(256), 0, c.stream, nf, (const double*)L.Z, (const double*)L.KZ,
                       (const double*)c.r, L.part);
    NCHK(hipGetLastError());
    NCHK(hipMemcpyAsync(L.h_part, L.part, sizeof(double) * NDOT * DOT_BLOCKS, hipMemcpyDeviceToHost, c.stream));
    NCHK(hipStreamSynchronize(c.stream));
    double d[NDOT];
    for (int k = 0; k < NDOT; ++k) {
        d[k] = 0.0;
        for (int b = 0; b < DOT_BLOCKS; ++b) d[k] += L.h_part[k * DOT_BLOCKS + b];
    }
    // Galerkin system G a = b over the usable columns (energy-optimal
    // combination: the guess never has a larger K-norm error than zero)
    double G[4][4], bb[4];
    {
        int k = 0;
        for (int i = 0; i < 4; ++i)
            for (int j = i; j < 4; ++j) G[i][j] = G[j][i] = d[k++];
        for (int i = 0; i < 4; ++i) bb[i] = d[10 + i];
    }
    bool act[4];
    double gmax = 0.0;
    for (int i = 0; i < 4; ++i) gmax = std::max(gmax, G[i][i]);
    for (int i = 0; i < 4; ++i) act[i] = gmax > 0.0 && G[i][i] > 1e-14 * gmax && (i < 3 ? use_net : have_prev);
    double alpha[4] = {0, 0, 0, 0};
    for (int attempt = 0; attempt < 4; ++attempt) {
        // Cholesky on the active set; a small pivot drops that column
        int idx[4], m = 0;
        for (int i = 0; i < 4; ++i)
            if (act[i]) idx[m++] = i;
        if (m == 0) break;
        double Lm[4][4] = {}, y[4];
        int drop = -1;
        for (int i = 0; i < m && drop < 0; ++i) {
            for (int j = 0; j <= i; ++j) {
                double s = G[idx[i]][idx[j]];
                for (int k = 0; k < j; ++k) s -= Lm[i][k] * Lm[j][k];
                if (i == j) {
                    if (s <= 1e-10 * G[idx[i]][idx[i]]) { drop = idx[i]; break; }
                    Lm[i][i] = std::sqrt(s);
                } else {
                    Lm[i][j] = s / Lm[j][j];
                }
            }
        }
        if (drop >= 0) { act[drop] = false; continue; }
        for (int i = 0; i < m; ++i) {
            double s = bb[idx[i]];
            for (int k = 0; k < i; ++k) s -= Lm[i][k] * y[k];
            y[i] = s / Lm[i][i];
        }
        for (int i = m - 1; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < m; ++k) s -= Lm[k][i] * alpha[idx[k]];
            alpha[idx[i]] = s / Lm[i][i];
        }
        break;
    }
    int cols = 0;
    for (int i = 0; i < 4; ++i) cols += alpha[i] != 0.0;
    for (int i = 0; i < 4; ++i) L.last_alpha[i] = alpha[i];
    L.last_cols = cols;
    if (cols == 0) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return false;
    }
    hipLaunchKernelGGL(k_combine, dim3(gfor(nf)), dim3(256), 0, c.stream, nf, (const double*)L.Z, alpha[0], alpha[1], alpha[2],
                       alpha[3], c.guess);
    NCHK(hipEventRecord(e1, c.stream));
    NCHK(hipEventSynchronize(e1));
    float ms;
    NCHK(hipEventElapsedTime(&ms, e0, e1));
    L.last_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return true;
}

}  // namespace pgfd

// dataset sample from device features / targets
static void net_add_device_sample(pgf_net* net, const float* feat, const float* tgt, hipStream_t st) {
    bf16* dx = nalloc<bf16>((size_t)NVOX * IN_CH);
    float* dt = nalloc<float>((size_t)NVOX * OUT_CH);
    net->ds_x.push_back(dx);
    net->ds_t.push_back(dt);
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)(((int64_t)NVOX * IN_CH + 255) / 256)), dim3(256), 0, st,
                       (int64_t)NVOX * IN_CH, feat, dx);
    NCHK(hipMemcpyAsync(dt, tgt, sizeof(float) * NVOX * OUT_CH, hipMemcpyDeviceToDevice, st));
    NCHK(hipStreamSynchronize(st));
}

template <typename F>
static int cguard(pgfd::Ctx* c, F&& f) {
    try {
        return f();
    } catch (const std::string& s) {
        c->err = s;
        return PGF_EDEVICE;
    }
}

extern "C" {

PGF_API int pgf_learn_attach(pgf_ctx* h, pgf_net* net, const double* box, const float* indicator) {
    if (!h || !box) return PGF_EARG;
    pgfd::Ctx* c = pgfd::ctx_of(h);
    // the voxel features and the prolongation map one node per free dof;
    // periodic images share dofs at different positions
    if (c->image) {
        c->err = "the learned guess does not support periodic meshes";
        return PGF_EUNSUPPORTED;
    }
    return cguard(c, [&]() -> int {
        if (c->learn) pgfd::learn_free(c->learn);
        c->learn = nullptr;
        auto* L = new pgfd::LearnState();
        c->learn = L;
        L->net = net;
        for (int d = 0; d < 3; ++d) {
            L->box.lo[d] = box[d];
            L->box.ext[d] = box[3 + d] - box[d];
            if (!(L->box.ext[d] > 0.0)) {
                if (d < c->dim) throw std::string("learn box must have positive extent");
                L->box.ext[d] = 1.0;
            }
        }
        L->ind = nalloc<float>(NVOX);
        if (indicator) NCHK(hipMemcpy(L->ind, indicator, sizeof(float) * NVOX, hipMemcpyHostToDevice));
        else NCHK(hipMemset(L->ind, 0, sizeof(float) * NVOX));
        L->feat = nalloc<float>((size_t)NVOX * IN_CH);
        L->cnt = nalloc<float>(NVOX);
        L->tgt = nalloc<float>((size_t)NVOX * OUT_CH);
        L->node_vox = nalloc<int32_t>(c->n_all);
        L->cpt_vox = nalloc<int32_t>(std::max<int64_t>(c->n_cpt, 1));
        L->U_prev = nalloc<double>((size_t)c->dim * c->n_all);
        L->U_prevprev = nalloc<double>((size_t)c->dim * c->n_all);
        L->U_start = nalloc<double>((size_t)c->dim * c->n_all);
        L->U_iter = nalloc<double>((size_t)c->dim * c->n_all);
        L->Z = nalloc<double>((size_t)4 * c->n_free);
        L->KZ = nalloc<double>((size_t)4 * c->n_free);
        L->part = nalloc<double>((size_t)pgfn::NDOT * pgfd::DOT_BLOCKS);
        NCHK(hipHostMalloc((void**)&L->h_part, sizeof(double) * pgfn::NDOT * pgfd::DOT_BLOCKS, hipHostMallocDefault));
        hipLaunchKernelGGL(pgfn::k_node_vox, dim3(pgfd::gfor(c->n_all)), dim3(256), 0, c->stream, c->n_all, (const double*)c->X,
                           L->box, L->node_vox);
        if (c->n_cut > 0)
            hipLaunchKernelGGL(pgfn::k_cpt_vox, dim3(pgfd::gfor(c->n_cut)), dim3(256), 0, c->stream, c->n_cut, c->dim, c->n_all,
                               (const double*)c->X, (const int32_t*)c->ce_nodes, (const int64_t*)c->cpt_ptr, L->box,
                               L->cpt_vox);
        NCHK(hipGetLastError());
        // the state at attach time is the first "previous" converged field
        NCHK(hipMemcpyAsync(L->U_prev, c->U, sizeof(double) * c->dim * c->n_all, hipMemcpyDeviceToDevice, c->stream));
        L->n_commits = 1;
        NCHK(hipStreamSynchronize(c->stream));
        return 0;
    });
}

PGF_API int pgf_learn_detach(pgf_ctx* h) {
    if (!h) return PGF_EARG;
    pgfd::Ctx* c = pgfd::ctx_of(h);
    if (c->learn) pgfd::learn_free(c->learn);
    c->learn = nullptr;
    return 0;
}

// the converged U of an accepted load step (the previous increment channel)
PGF_API int pgf_learn_commit(pgf_ctx* h) {
    if (!h) return PGF_EARG;
    pgfd::Ctx* c = pgfd::ctx_of(h);
    if (!c->learn) return PGF_ESTATE;
    return cguard(c, [&]() -> int {
        pgfd::LearnState& L = *c->learn;
        const size_t n = sizeof(double) * c->dim * c->n_all;
        NCHK(hipMemcpyAsync(L.U_prevprev, L.U_prev, n, hipMemcpyDeviceToDevice, c->stream));
        NCHK(hipMemcpyAsync(L.U_prev, c->U, n, hipMemcpyDeviceToDevice, c->stream));
        NCHK(hipStreamSynchronize(c->stream));
        ++L.n_commits;
        return 0;
    });
}

// training sample from the last step: its start features and the normalised
// voxel average of the converged increment U - U_start
PGF_API int pgf_learn_record(pgf_ctx* h) {
    if (!h) return PGF_EARG;
    pgfd::Ctx* c = pgfd::ctx_of(h);
    if (!c->learn || !c->learn->net || !c->learn->have_features) return PGF_ESTATE;
    return cguard(c, [&]() -> int {
        pgfd::LearnState& L = *c->learn;
        NCHK(hipMemsetAsync(L.tgt, 0, sizeof(float) * NVOX * OUT_CH, c->stream));
        NCHK(hipMemsetAsync(L.cnt, 0, sizeof(float) * NVOX, c->stream));
        hipLaunchKernelGGL(pgfn::k_feat_nodal, dim3(pgfd::gfor(c->n_all)), dim3(256), 0, c->stream, c->n_all, c->dim,
                           (const double*)c->U, (const double*)L.U_start, (const int32_t*)L.node_vox, L.tgt, OUT_CH, 0, L.cnt);
        hipLaunchKernelGGL(pgfn::k_feat_avg, dim3(pgfd::gfor(NVOX)), dim3(256), 0, c->stream, L.tgt, OUT_CH, 0, c->dim,
                           (const float*)L.cnt);
        NCHK(hipGetLastError());
        pgfd::normalise(*c, L, L.tgt, OUT_CH, 0, 3);
        net_add_device_sample(L.net, L.feat, L.tgt, c->stream);
        return 0;
    });
}

// last guess: out[0..3] Galerkin coefficients (3 network channels, previous
// increment), out[4] number of columns used, out[5] device time of the guess (ms)
PGF_API int pgf_learn_last(pgf_ctx* h, double* out) {
    if (!h || !out) return PGF_EARG;
    pgfd::Ctx* c = pgfd::ctx_of(h);
    if (!c->learn) return PGF_ESTATE;
    for (int i = 0; i < 4; ++i) out[i] = c->learn->last_alpha[i];
    out[4] = c->learn->last_cols;
    out[5] = c->learn->last_ms;
    return 0;
}

// host copy of the last step-start features [64^3][8]
PGF_API int pgf_learn_features(pgf_ctx* h, float* out) {
    if (!h || !out) return PGF_EARG;
    pgfd::Ctx* c = pgfd::ctx_of(h);
    if (!c->learn) return PGF_ESTATE;
    return cguard(c, [&]() -> int {
        NCHK(hipMemcpy(out, c->learn->feat, sizeof(float) * NVOX * IN_CH, hipMemcpyDeviceToHost));
        return 0;
    });
}

}  // extern "C"
