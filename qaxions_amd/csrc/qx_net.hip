// qx_net.hip -- learned coarse-step corrector: four periodic 3x3 convolutions
// 3 -> 32 -> 32 -> 32 -> 2 (GELU after the first three) applied to
// (Re psi, Im psi, a V) of a 2D lattice, psi <- psi + s net(psi, a V).  No
// reference counterpart.
//
// Why V is an input: the splitting error of a KDK step is a sum of nested
// commutators of T = -laplacian/2 and V, with terms in grad V and the
// Hessian of V.  V is a nonlocal functional of psi (a Poisson solve), so a
// network with a 9 x 9 receptive field cannot form those terms from psi
// alone; given V it can (they are local in V).  The lattice always holds the
// V of the state being corrected.
//
// Numerics: weights are held in FP32 (master copy, Adam state) and rounded
// to BF16 for every forward/backward use; the hidden activations are stored
// in BF16; all accumulation is FP32.  The 32 -> 32 layers of the forward
// pass run on MFMA (v_mfma_f32_16x16x16_bf16): output pixels x channels =
// (shifted input patches) x (tap weights), K = 9 taps x 32 channels; the
// thin layers and the backward pass are FP32 vector kernels.
//
// Layouts: activations [P][C] (channel fastest, P = H*W row-major); the
// input is [P][3] (Re psi, Im psi, a V), the output a complex64 psi array
// [P][2]; kernel weights
// [tap][ci][co] with tap = 3*dy + dx reading x[(h+dy-1) mod H][(w+dx-1) mod W];
// the flat parameter vector is, per layer, W[co][ci][dy][dx] then b[co].
#include "../../include/qx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace qx {
qx_status leapfrog_enqueue(qx_ctx *c, double ds, int nsteps, int first_half, int last_half);
qx_status update_potential_enqueue(qx_ctx *c);
unsigned *drift_tick(qx_ctx *c);
}  // namespace qx

namespace {

thread_local std::string g_net_err;
qx_status nfail(qx_status s, const std::string &m)
{
    g_net_err = m;
    return s;
}
#define NCHK(x)                                                                            \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return nfail(QX_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define NTRY(x)                         \
    do {                                \
        qx_status s_ = (x);             \
        if (s_ != QX_OK) return s_;     \
    } while (0)

constexpr int NL = 4;
constexpr int CH[NL + 1] = {3, 32, 32, 32, 2};
constexpr int NPARAM = 19970;
constexpr int CIN = 3;  // input channels

__host__ __device__ constexpr int w_count(int l) { return CH[l] * CH[l + 1] * 9; }
__host__ __device__ constexpr int layer_off(int l) { return l == 0 ? 0 : layer_off(l - 1) + w_count(l - 1) + CH[l]; }

typedef unsigned short bf16_t;

__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ bf16_t f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite values)
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z)
{
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

// ---------------------------------------------------------------------------
// parameter packing: flat FP32 theta -> BF16 kernel weights [tap][ci][co]
// (forward) and [tap'][co][ci] with tap' = 8 - tap (transposed, for the
// input gradient), biases stay FP32
// ---------------------------------------------------------------------------
__global__ void pack_kernel(const float *__restrict__ theta, bf16_t *__restrict__ wf,
                            bf16_t *__restrict__ wt, int ci, int co, int toff, int woff)
{
    const int n = ci * co * 9;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        // canonical index i = ((o*ci + c)*3 + dy)*3 + dx
        const int tap = i % 9, c = (i / 9) % ci, o = i / (9 * ci);
        const bf16_t b = f2bf(theta[toff + i]);
        wf[woff + (tap * ci + c) * co + o] = b;
        wt[woff + ((8 - tap) * co + o) * ci + c] = b;
    }
}

// ---------------------------------------------------------------------------
// vector convolution: one thread per pixel, all CO outputs
//   IN = float (first layer, psi) or bf16 (hidden activations)
//   MODE 0: forward, z = conv + b, a = gelu(z) -> bf16 (hidden layers)
//   MODE 1: forward last layer: out = x[:, 0:2] + oscale (conv + b), float
//           [P][2]; x (the network input, x_res) has CIN channels per pixel
//   MODE 2: input gradient of a hidden layer: d_in[p][c] = conv_T(dz) then
//           multiplied by gelu'(z_prev[p][c]) (the producing layer's pre-act)
//   MODE 3: input gradient without the GELU factor (gradient w.r.t. psi)
// ---------------------------------------------------------------------------
template <int CI, int CO, int MODE, typename IN>
__global__ void __launch_bounds__(256)
conv_vec(const IN *__restrict__ in, int H, int W, const bf16_t *__restrict__ wk,
         const float *__restrict__ bias, float *__restrict__ z_out, void *__restrict__ out,
         const float *__restrict__ x_res, const float *__restrict__ zprev, float oscale = 1.0f)
{
    __shared__ float ws[9 * CI * CO];
    for (int i = threadIdx.x; i < 9 * CI * CO; i += blockDim.x) ws[i] = bf2f(wk[i]);
    __syncthreads();
    const long long P = (long long)H * W;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int h = (int)(p / W), w = (int)(p % W);
    float acc[CO];
#pragma unroll
    for (int o = 0; o < CO; ++o) acc[o] = (MODE <= 1) ? bias[o] : 0.0f;
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3, dx = tap % 3;
        int hh = h + dy - 1, ww = w + dx - 1;
        hh += (hh < 0) ? H : 0;
        hh -= (hh >= H) ? H : 0;
        ww += (ww < 0) ? W : 0;
        ww -= (ww >= W) ? W : 0;
        const IN *src = in + ((long long)hh * W + ww) * CI;
        const float *wt = ws + tap * CI * CO;
        // the pixel's CI inputs in 16-byte loads
        constexpr int VEC = 16 / (int)sizeof(IN);
        static_assert(CI % VEC == 0 || CI < VEC, "channel count must fill 16-byte loads");
        if constexpr (CI >= VEC) {
#pragma unroll
            for (int c0 = 0; c0 < CI; c0 += VEC) {
                const uint4 raw = *(const uint4 *)(src + c0);
                const IN *xv = (const IN *)&raw;
#pragma unroll
                for (int cc = 0; cc < VEC; ++cc) {
                    float x;
                    if constexpr (sizeof(IN) == 2) x = bf2f(xv[cc]);
                    else x = (float)xv[cc];
#pragma unroll
                    for (int o = 0; o < CO; ++o) acc[o] = fmaf(wt[(c0 + cc) * CO + o], x, acc[o]);
                }
            }
        } else {
#pragma unroll
            for (int c = 0; c < CI; ++c) {
                float x;
                if constexpr (sizeof(IN) == 2) x = bf2f(src[c]);
                else x = (float)src[c];
#pragma unroll
                for (int o = 0; o < CO; ++o) acc[o] = fmaf(wt[c * CO + o], x, acc[o]);
            }
        }
    }
    if constexpr (MODE == 0) {
        // packed 16-byte stores of the pixel's CO outputs
        static_assert(CO % 8 == 0, "hidden layers have a multiple of 8 channels");
        if (z_out) {
#pragma unroll
            for (int o = 0; o < CO; o += 4)
                *(float4 *)(z_out + p * CO + o) = float4{acc[o], acc[o + 1], acc[o + 2], acc[o + 3]};
        }
        bf16_t *a = (bf16_t *)out;
#pragma unroll
        for (int o = 0; o < CO; o += 8) {
            uint4 pk;
            pk.x = (unsigned)f2bf(gelu(acc[o])) | ((unsigned)f2bf(gelu(acc[o + 1])) << 16);
            pk.y = (unsigned)f2bf(gelu(acc[o + 2])) | ((unsigned)f2bf(gelu(acc[o + 3])) << 16);
            pk.z = (unsigned)f2bf(gelu(acc[o + 4])) | ((unsigned)f2bf(gelu(acc[o + 5])) << 16);
            pk.w = (unsigned)f2bf(gelu(acc[o + 6])) | ((unsigned)f2bf(gelu(acc[o + 7])) << 16);
            *(uint4 *)(a + p * CO + o) = pk;
        }
    } else if constexpr (MODE == 1) {
        float *y = (float *)out;
#pragma unroll
        for (int o = 0; o < CO; ++o) y[p * CO + o] = fmaf(oscale, acc[o], x_res[p * CIN + o]);
    } else if constexpr (MODE == 2) {
        float *d = (float *)out;
#pragma unroll
        for (int o = 0; o < CO; o += 4) {
            const float4 zp = *(const float4 *)(zprev + p * CO + o);
            *(float4 *)(d + p * CO + o) = float4{acc[o] * gelu_grad(zp.x), acc[o + 1] * gelu_grad(zp.y),
                                                 acc[o + 2] * gelu_grad(zp.z), acc[o + 3] * gelu_grad(zp.w)};
        }
    } else {
        float *d = (float *)out;
#pragma unroll
        for (int o = 0; o < CO; ++o) d[p * CO + o] = acc[o];
    }
}

// ---------------------------------------------------------------------------
// MFMA forward of a 32 -> 32 layer.  A wave computes 16 consecutive pixels of
// one row x 32 output channels per tile: two 16x16 f32 accumulators over 18
// K-chunks (9 taps x 2 halves of 16 input channels).  Weight fragments stay
// in registers for all tiles of the wave.
//   A (16 px x 16 ch):  lane l holds pixel l%16, channels 4*(l/16)+j
//   B (16 ch x 16 co):  lane l holds out channel l%16, in channels 4*(l/16)+j
//   D (16 px x 16 co):  lane l holds out channel l%16, pixels 4*(l/16)+j
// ---------------------------------------------------------------------------
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

template <bool KEEP_Z>
__global__ void __launch_bounds__(256)
conv32_mfma(const bf16_t *__restrict__ in, int H, int W, const bf16_t *__restrict__ wk,
            const float *__restrict__ bias, float *__restrict__ z_out, bf16_t *__restrict__ a_out,
            int tiles_per_wave)
{
    // per-wave output tile [16 px][32 ch] staged in LDS so that the wave
    // writes 16 contiguous pixel records with 16-byte stores
    __shared__ bf16_t otile[4][16 * 32];
    __shared__ float ztile[KEEP_Z ? 4 : 1][KEEP_Z ? 16 * 32 : 1];
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    // weight fragments: [tap][half][ntile]
    short4_t bfr[9][2][2];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int hf = 0; hf < 2; ++hf)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                short4_t v;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    v[j] = (short)wk[(tap * 32 + hf * 16 + 4 * g + j) * 32 + nt * 16 + r];
                bfr[tap][hf][nt] = v;
            }
    const float b0 = bias[r], b1 = bias[16 + r];
    const int tiles_per_row = W / 16;
    const long long ntiles = (long long)H * tiles_per_row;
    for (int it = 0; it < tiles_per_wave; ++it) {
        const long long tile = (long long)wave * tiles_per_wave + it;
        if (tile >= ntiles) break;
        const int h = (int)(tile / tiles_per_row), w0 = (int)(tile % tiles_per_row) * 16;
        float4_t acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;
            int hh = h + dy - 1, ww = w0 + r + dx - 1;
            hh += (hh < 0) ? H : 0;
            hh -= (hh >= H) ? H : 0;
            ww += (ww < 0) ? W : 0;
            ww -= (ww >= W) ? W : 0;
            const bf16_t *src = in + ((long long)hh * W + ww) * 32 + 4 * g;
#pragma unroll
            for (int hf = 0; hf < 2; ++hf) {
                const short4_t a = *(const short4_t *)(src + hf * 16);
                acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bfr[tap][hf][0], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bfr[tap][hf][1], acc1, 0, 0, 0);
            }
        }
        const int wv = threadIdx.x >> 6;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int px = 4 * g + j;
            const float z0 = acc0[j] + b0, z1 = acc1[j] + b1;
            if (KEEP_Z) {
                ztile[wv][px * 32 + r] = z0;
                ztile[wv][px * 32 + 16 + r] = z1;
            }
            otile[wv][px * 32 + r] = f2bf(gelu(z0));
            otile[wv][px * 32 + 16 + r] = f2bf(gelu(z1));
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's LDS writes are done
        const long long p0 = (long long)h * W + w0;
        // 16 px x 32 ch bf16 = 1 KiB: 64 lanes x 16 B
        *(uint4 *)(a_out + p0 * 32 + lane * 8) = *(const uint4 *)(&otile[wv][lane * 8]);
        if (KEEP_Z) {
            // 16 px x 32 ch f32 = 2 KiB: 64 lanes x 2 x 16 B
            *(float4 *)(z_out + p0 * 32 + lane * 4) = *(const float4 *)(&ztile[wv][lane * 4]);
            *(float4 *)(z_out + p0 * 32 + 256 + lane * 4) = *(const float4 *)(&ztile[wv][256 + lane * 4]);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------------------
// weight gradient: gw[tap][ci][co] = sum_p in[p + tap][ci] dz[p][co],
// gb[co] = sum_p dz[p][co].  Block b reduces pixels [b*CHUNK, (b+1)*CHUNK)
// into a partial row; partials are summed in a fixed order afterwards.
// ---------------------------------------------------------------------------
template <int CI, int CO, typename IN>
__global__ void __launch_bounds__(256)
wgrad_kernel(const IN *__restrict__ in, const float *__restrict__ dz, int H, int W, int chunk,
             float *__restrict__ partial)
{
    constexpr int NW = 9 * CI * CO;
    constexpr int PER = (NW + CO + 255) / 256;
    constexpr int TP = 32;  // pixels staged per round
    __shared__ float xs[TP][9][CI];
    __shared__ float ds[TP][CO];
    float acc[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) acc[k] = 0.f;
    const long long P = (long long)H * W;
    const long long p0 = (long long)blockIdx.x * chunk;
    for (long long pb = p0; pb < p0 + chunk && pb < P; pb += TP) {
        __syncthreads();
        for (int i = threadIdx.x; i < TP * 9 * CI; i += blockDim.x) {
            const int c = i % CI, tap = (i / CI) % 9, q = i / (9 * CI);
            const long long p = pb + q;
            float v = 0.f;
            if (p < P && p < p0 + chunk) {
                const int h = (int)(p / W), w = (int)(p % W);
                int hh = h + tap / 3 - 1, ww = w + tap % 3 - 1;
                hh += (hh < 0) ? H : 0;
                hh -= (hh >= H) ? H : 0;
                ww += (ww < 0) ? W : 0;
                ww -= (ww >= W) ? W : 0;
                const IN x = in[((long long)hh * W + ww) * CI + c];
                if constexpr (sizeof(IN) == 2) v = bf2f(x);
                else v = (float)x;
            }
            xs[q][tap][c] = v;
        }
        for (int i = threadIdx.x; i < TP * CO; i += blockDim.x) {
            const int o = i % CO, q = i / CO;
            const long long p = pb + q;
            ds[q][o] = (p < P && p < p0 + chunk) ? dz[p * CO + o] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = threadIdx.x + 256 * k;
            if (e < NW) {
                const int o = e % CO, c = (e / CO) % CI, tap = e / (CO * CI);
                float s = acc[k];
                for (int q = 0; q < TP; ++q) s = fmaf(xs[q][tap][c], ds[q][o], s);
                acc[k] = s;
            } else if (e < NW + CO) {
                const int o = e - NW;
                float s = acc[k];
                for (int q = 0; q < TP; ++q) s += ds[q][o];
                acc[k] = s;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = threadIdx.x + 256 * k;
        if (e < NW + CO) partial[(long long)blockIdx.x * (NW + CO) + e] = acc[k];
    }
}

// sum partial rows (fixed order) into the flat FP32 gradient at layer offset
__global__ void wgrad_reduce(const float *__restrict__ partial, int nblocks, int ci, int co,
                             int toff, float *__restrict__ grad)
{
    const int NW = 9 * ci * co;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < NW + co; e += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < nblocks; ++b) s += partial[(long long)b * (NW + co) + e];
        int dst;
        if (e < NW) {
            const int o = e % co, c = (e / co) % ci, tap = e / (co * ci);
            dst = toff + (o * ci + c) * 9 + tap;  // canonical W[co][ci][dy][dx]
        } else {
            dst = toff + NW + (e - NW);
        }
        grad[dst] = (float)s;
    }
}

// residual r = y - target; d = oscale 2 r / n is the gradient with respect
// to the last layer's output; per-block partial of sum r^2
__global__ void __launch_bounds__(256)
loss_kernel(const float *__restrict__ y, const float *__restrict__ t, long long n, float oscale,
            float *__restrict__ d, double *__restrict__ partial)
{
    __shared__ double sh[4];
    double s = 0.0;
    const float sc = (float)(2.0 * (double)oscale / (double)n);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const float r = y[i] - t[i];
        d[i] = sc * r;
        s += (double)r * r;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void adam_kernel(float *__restrict__ theta, float *__restrict__ m, float *__restrict__ v,
                            const float *__restrict__ g, int n, float lr, float b1, float b2, float eps,
                            float c1, float c2)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        theta[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// network input from a lattice: (Re psi, Im psi, vscale V)
__global__ void input_kernel(const float *__restrict__ psi, const float *__restrict__ V, float vscale,
                             long long P, float *__restrict__ x)
{
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P;
         p += (long long)gridDim.x * blockDim.x) {
        x[3 * p] = psi[2 * p];
        x[3 * p + 1] = psi[2 * p + 1];
        x[3 * p + 2] = vscale * V[p];
    }
}

// sum |y|^2 of P complex values: FP64 per-block partials (grid <= 1024)
__global__ void __launch_bounds__(256)
norm2_partial(const float *__restrict__ y, long long P, double *__restrict__ partial)
{
    __shared__ double sh[4];
    double s = 0.0;
    for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < P;
         p += (long long)gridDim.x * blockDim.x)
        s += (double)y[2 * p] * y[2 * p] + (double)y[2 * p + 1] * y[2 * p + 1];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}
// psi = y sqrt(target / sum of the partials): the corrected state scaled
// back to the mass it had before the correction (partials summed in order
// by every thread, an FP64 sum of at most 1024 terms)
__global__ void __launch_bounds__(256)
renorm_copy(const float *__restrict__ y, const double *__restrict__ partial, int nb, double target,
            long long n, float *__restrict__ psi)
{
    __shared__ float scale;
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += partial[b];
        scale = (float)sqrt(target / s);
    }
    __syncthreads();
    const float f = scale;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        psi[i] = y[i] * f;
}

unsigned grid1(long long n, int t = 256)
{
    long long b = (n + t - 1) / t;
    if (b > 4096) b = 4096;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

struct qx_net {
    int H = 0, W = 0, device = 0;
    long long P = 0;
    hipStream_t stream = nullptr;
    float *theta = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    bf16_t *wf = nullptr, *wt = nullptr;  // packed weights (forward, transposed)
    float *z[3] = {nullptr, nullptr, nullptr};  // pre-activations of the hidden layers
    bf16_t *a[3] = {nullptr, nullptr, nullptr}; // hidden activations
    float *y = nullptr;      // network output x + net(x) [P][2]
    float *d0 = nullptr, *d1 = nullptr;  // gradient buffers [P][32]
    float *partial = nullptr;
    double *lpart = nullptr, *h_lpart = nullptr;
    float *xbuf = nullptr, *tbuf = nullptr;  // host-array staging ([P][3], [P][2])
    float *xin = nullptr;    // input built from a lattice [P][3]
    int adam_t = 0;
    float out_scale = 1.0f;  // y = x + out_scale * net(x)
    float in_scale = 1.0f;   // V channel = in_scale * V for lattice inputs
    bool packed = false;
    int wchunk = 512, wblocks = 0;
    void *save_psi = nullptr, *save_V = nullptr;  // fallback copies (lazy)
    unsigned *save_tick = nullptr;                // and of the lattice's drift counter
    double energy_tol = 0.0;  // coarse_steps: energy-drift gate per chunk (0: off)
};

namespace {

int wofs(int l)  // offset of layer l's packed weights
{
    int o = 0;
    for (int k = 0; k < l; ++k) o += w_count(k);
    return o;
}

qx_status pack(qx_net *n)
{
    for (int l = 0; l < NL; ++l)
        pack_kernel<<<grid1(w_count(l)), 256, 0, n->stream>>>(n->theta, n->wf, n->wt, CH[l], CH[l + 1],
                                                             layer_off(l), wofs(l));
    n->packed = true;
    NCHK(hipGetLastError());
    return QX_OK;
}

const float *bias(const qx_net *n, int l) { return n->theta + layer_off(l) + w_count(l); }

// forward pass: x [P][3] float -> y [P][2] = x[:, 0:2] + s net(x); keep z
// for training
qx_status forward(qx_net *n, const float *x, float *y, bool keep, hipStream_t s = nullptr)
{
    if (!s) s = n->stream;
    if (!n->packed) NTRY(pack(n));
    const unsigned gp = (unsigned)((n->P + 255) / 256);
    conv_vec<CIN, 32, 0, float><<<gp, 256, 0, s>>>(x, n->H, n->W, n->wf + wofs(0), bias(n, 0),
                                                 keep ? n->z[0] : nullptr, n->a[0], nullptr, nullptr);
    const bool mfma = (n->W % 16) == 0;
    for (int l = 1; l <= 2; ++l) {
        if (mfma) {
            const long long ntiles = n->P / 16;
            const int tpw = 4;
            const long long waves = (ntiles + tpw - 1) / tpw;
            const unsigned blocks = (unsigned)((waves + 3) / 4);
            if (keep)
                conv32_mfma<true><<<blocks, 256, 0, s>>>(n->a[l - 1], n->H, n->W, n->wf + wofs(l), bias(n, l),
                                                         n->z[l], n->a[l], tpw);
            else
                conv32_mfma<false><<<blocks, 256, 0, s>>>(n->a[l - 1], n->H, n->W, n->wf + wofs(l), bias(n, l),
                                                          nullptr, n->a[l], tpw);
        } else {
            conv_vec<32, 32, 0, bf16_t><<<gp, 256, 0, s>>>(n->a[l - 1], n->H, n->W, n->wf + wofs(l), bias(n, l),
                                                           keep ? n->z[l] : nullptr, n->a[l], nullptr, nullptr);
        }
    }
    conv_vec<32, 2, 1, bf16_t><<<gp, 256, 0, s>>>(n->a[2], n->H, n->W, n->wf + wofs(3), bias(n, 3), nullptr,
                                                  y, x, nullptr, n->out_scale);
    NCHK(hipGetLastError());
    return QX_OK;
}

qx_status nsync(qx_net *n)
{
    NCHK(hipGetLastError());
    NCHK(hipStreamSynchronize(n->stream));
    return QX_OK;
}

// loss and full gradient into n->grad; x device [P][3], t device [P][2]
qx_status loss_grad(qx_net *n, const float *x, const float *t, double *loss)
{
    NTRY(forward(n, x, n->y, true));
    hipStream_t s = n->stream;
    const long long ne = n->P * 2;
    const unsigned gl = grid1(ne);
    loss_kernel<<<gl, 256, 0, s>>>(n->y, t, ne, n->out_scale, n->d0, n->lpart);
    // layer 3 (32 -> 2, no activation): dz3 = d0 [P][2]
    const unsigned gp = (unsigned)((n->P + 255) / 256);
    const int wb = n->wblocks;
    wgrad_kernel<32, 2, bf16_t><<<wb, 256, 0, s>>>(n->a[2], n->d0, n->H, n->W, n->wchunk, n->partial);
    wgrad_reduce<<<grid1(9 * 32 * 2 + 2), 256, 0, s>>>(n->partial, wb, 32, 2, layer_off(3), n->grad);
    // d a2 -> dz2 = (W3^T dz3) * gelu'(z2)
    conv_vec<2, 32, 2, float><<<gp, 256, 0, s>>>(n->d0, n->H, n->W, n->wt + wofs(3), nullptr, nullptr, n->d1,
                                                 nullptr, n->z[2]);
    for (int l = 2; l >= 1; --l) {
        float *dz = (l == 2) ? n->d1 : n->d0;
        float *dn = (l == 2) ? n->d0 : n->d1;
        wgrad_kernel<32, 32, bf16_t><<<wb, 256, 0, s>>>(n->a[l - 1], dz, n->H, n->W, n->wchunk, n->partial);
        wgrad_reduce<<<grid1(9 * 32 * 32 + 32), 256, 0, s>>>(n->partial, wb, 32, 32, layer_off(l), n->grad);
        conv_vec<32, 32, 2, float><<<gp, 256, 0, s>>>(dz, n->H, n->W, n->wt + wofs(l), nullptr, nullptr, dn,
                                                      nullptr, n->z[l - 1]);
    }
    // layer 0 (3 -> 32): dz0 in d1
    wgrad_kernel<CIN, 32, float><<<wb, 256, 0, s>>>(x, n->d1, n->H, n->W, n->wchunk, n->partial);
    wgrad_reduce<<<grid1(9 * CIN * 32 + 32), 256, 0, s>>>(n->partial, wb, CIN, 32, layer_off(0), n->grad);
    if (loss) {
        NCHK(hipMemcpyAsync(n->h_lpart, n->lpart, sizeof(double) * gl, hipMemcpyDeviceToHost, s));
        NTRY(nsync(n));
        double sum = 0.0;
        for (unsigned b = 0; b < gl; ++b) sum += n->h_lpart[b];
        *loss = sum / (double)ne;
    }
    return QX_OK;
}

}  // namespace

extern "C" {

const char *qx_net_last_error(void) { return g_net_err.c_str(); }
int qx_net_param_count(void) { return NPARAM; }

qx_status qx_net_create(int H, int W, int device, qx_net **out)
{
    if (!out) return nfail(QX_EINVAL, "out is null");
    *out = nullptr;
    if (H < 3 || W < 3) return nfail(QX_EINVAL, "lattice must be at least 3x3");
    NCHK(hipSetDevice(device));
    std::unique_ptr<qx_net> n(new qx_net());
    n->H = H;
    n->W = W;
    n->device = device;
    n->P = (long long)H * W;
    n->wblocks = (int)((n->P + n->wchunk - 1) / n->wchunk);
    auto cleanup = [&](qx_status s) {
        qx_net_destroy(n.release());
        return s;
    };
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(nfail(QX_EDEVICE, "stream creation failed"));
    const int NWT = wofs(NL);
    struct A { void **p; size_t b; } al[] = {
        {(void **)&n->theta, sizeof(float) * NPARAM}, {(void **)&n->grad, sizeof(float) * NPARAM},
        {(void **)&n->m, sizeof(float) * NPARAM}, {(void **)&n->v, sizeof(float) * NPARAM},
        {(void **)&n->wf, sizeof(bf16_t) * NWT}, {(void **)&n->wt, sizeof(bf16_t) * NWT},
        {(void **)&n->z[0], sizeof(float) * n->P * 32}, {(void **)&n->z[1], sizeof(float) * n->P * 32},
        {(void **)&n->z[2], sizeof(float) * n->P * 32}, {(void **)&n->a[0], sizeof(bf16_t) * n->P * 32},
        {(void **)&n->a[1], sizeof(bf16_t) * n->P * 32}, {(void **)&n->a[2], sizeof(bf16_t) * n->P * 32},
        {(void **)&n->y, sizeof(float) * n->P * 2}, {(void **)&n->d0, sizeof(float) * n->P * 32},
        {(void **)&n->d1, sizeof(float) * n->P * 32},
        {(void **)&n->partial, sizeof(float) * (size_t)n->wblocks * (9 * 32 * 32 + 32)},
        {(void **)&n->lpart, sizeof(double) * 4096}, {(void **)&n->xbuf, sizeof(float) * n->P * CIN},
        {(void **)&n->tbuf, sizeof(float) * n->P * 2}, {(void **)&n->xin, sizeof(float) * n->P * CIN}};
    for (auto &x : al)
        if (hipMalloc(x.p, x.b) != hipSuccess) {
            *x.p = nullptr;
            return cleanup(nfail(QX_ENOMEM, "network allocation failed"));
        }
    if (hipHostMalloc((void **)&n->h_lpart, sizeof(double) * 4096) != hipSuccess)
        return cleanup(nfail(QX_ENOMEM, "pinned allocation failed"));
    if (hipMemset(n->theta, 0, sizeof(float) * NPARAM) != hipSuccess ||
        hipMemset(n->m, 0, sizeof(float) * NPARAM) != hipSuccess ||
        hipMemset(n->v, 0, sizeof(float) * NPARAM) != hipSuccess ||
        hipMemset(n->grad, 0, sizeof(float) * NPARAM) != hipSuccess)
        return cleanup(nfail(QX_EDEVICE, "initialisation failed"));
    *out = n.release();
    return QX_OK;
}

qx_status qx_net_destroy(qx_net *n)
{
    if (!n) return QX_OK;
    (void)hipSetDevice(n->device);
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    void *b[] = {n->theta, n->grad, n->m, n->v, n->wf, n->wt, n->z[0], n->z[1], n->z[2], n->a[0], n->a[1],
                 n->a[2], n->y, n->d0, n->d1, n->partial, n->lpart, n->xbuf, n->tbuf, n->xin, n->save_psi,
                 n->save_V, n->save_tick};
    for (void *p : b)
        if (p) (void)hipFree(p);
    if (n->h_lpart) (void)hipHostFree(n->h_lpart);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
    return QX_OK;
}

qx_status qx_net_set_params(qx_net *n, const float *theta)
{
    if (!n || !theta) return nfail(QX_EINVAL, "null argument");
    NCHK(hipSetDevice(n->device));
    NCHK(hipMemcpy(n->theta, theta, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    NCHK(hipMemset(n->m, 0, sizeof(float) * NPARAM));
    NCHK(hipMemset(n->v, 0, sizeof(float) * NPARAM));
    n->adam_t = 0;
    NTRY(pack(n));
    return nsync(n);
}

qx_status qx_net_get_params(qx_net *n, float *theta)
{
    if (!n || !theta) return nfail(QX_EINVAL, "null argument");
    NCHK(hipSetDevice(n->device));
    NCHK(hipMemcpy(theta, n->theta, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    return QX_OK;
}

qx_status qx_net_forward(qx_net *n, const void *x_dev, void *y_dev)
{
    if (!n || !x_dev || !y_dev) return nfail(QX_EINVAL, "null argument");
    NCHK(hipSetDevice(n->device));
    NTRY(forward(n, (const float *)x_dev, (float *)y_dev, false));
    return nsync(n);
}

qx_status qx_net_forward_host(qx_net *n, const float *x, float *y)
{
    if (!n || !x || !y) return nfail(QX_EINVAL, "null argument");
    NCHK(hipSetDevice(n->device));
    NCHK(hipMemcpy(n->xbuf, x, sizeof(float) * n->P * CIN, hipMemcpyHostToDevice));
    NTRY(forward(n, n->xbuf, n->y, false));
    NTRY(nsync(n));
    NCHK(hipMemcpy(y, n->y, sizeof(float) * n->P * 2, hipMemcpyDeviceToHost));
    return QX_OK;
}

qx_status qx_net_loss_grad(qx_net *n, const void *x_dev, const void *t_dev, double *loss, float *grad_host)
{
    if (!n || !x_dev || !t_dev) return nfail(QX_EINVAL, "null argument");
    NCHK(hipSetDevice(n->device));
    NTRY(loss_grad(n, (const float *)x_dev, (const float *)t_dev, loss));
    NTRY(nsync(n));
    if (grad_host) NCHK(hipMemcpy(grad_host, n->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    return QX_OK;
}

qx_status qx_net_loss_grad_host(qx_net *n, const float *x, const float *t, double *loss, float *grad)
{
    if (!n || !x || !t) return nfail(QX_EINVAL, "null argument");
    NCHK(hipSetDevice(n->device));
    NCHK(hipMemcpy(n->xbuf, x, sizeof(float) * n->P * CIN, hipMemcpyHostToDevice));
    NCHK(hipMemcpy(n->tbuf, t, sizeof(float) * n->P * 2, hipMemcpyHostToDevice));
    return qx_net_loss_grad(n, n->xbuf, n->tbuf, loss, grad);
}

qx_status qx_net_adam_step(qx_net *n, float lr, float b1, float b2, float eps)
{
    if (!n) return nfail(QX_EINVAL, "null network");
    NCHK(hipSetDevice(n->device));
    n->adam_t += 1;
    const float c1 = 1.f - std::pow(b1, (float)n->adam_t), c2 = 1.f - std::pow(b2, (float)n->adam_t);
    adam_kernel<<<grid1(NPARAM), 256, 0, n->stream>>>(n->theta, n->m, n->v, n->grad, NPARAM, lr, b1, b2, eps,
                                                      c1, c2);
    NTRY(pack(n));
    return nsync(n);
}

// one training step on device pairs: loss, gradient, Adam
qx_status qx_net_train_step(qx_net *n, const void *x_dev, const void *t_dev, float lr, double *loss)
{
    if (!n || !x_dev || !t_dev) return nfail(QX_EINVAL, "null argument");
    NCHK(hipSetDevice(n->device));
    NTRY(loss_grad(n, (const float *)x_dev, (const float *)t_dev, loss));
    return qx_net_adam_step(n, lr, 0.9f, 0.999f, 1e-8f);
}

void *qx_net_stream(qx_net *n) { return n ? (void *)n->stream : nullptr; }

static qx_status check_ctx(qx_net *n, qx_ctx *ctx, float **psi, float **V);
static void lattice_input(qx_net *n, const float *psi, const float *V, hipStream_t s);

// one training step on a pair of lattices: input (psi, V) of `coarse`,
// target psi of `fine`
qx_status qx_net_train_step_lattices(qx_net *n, qx_ctx *coarse, qx_ctx *fine, float lr, double *loss)
{
    float *pc, *vc, *pf, *vf;
    NTRY(check_ctx(n, coarse, &pc, &vc));
    NTRY(check_ctx(n, fine, &pf, &vf));
    NCHK(hipSetDevice(n->device));
    NCHK(hipStreamSynchronize((hipStream_t)qx_stream(coarse)));
    NCHK(hipStreamSynchronize((hipStream_t)qx_stream(fine)));
    lattice_input(n, pc, vc, n->stream);
    NTRY(loss_grad(n, n->xin, pf, loss));
    return qx_net_adam_step(n, lr, 0.9f, 0.999f, 1e-8f);
}

// mean((x - t)^2) over a pair: the loss of an identity corrector
qx_status qx_net_pair_mse(qx_net *n, const void *x_dev, const void *t_dev, double *mse)
{
    if (!n || !x_dev || !t_dev || !mse) return nfail(QX_EINVAL, "null argument");
    NCHK(hipSetDevice(n->device));
    const long long ne = n->P * 2;
    const unsigned gl = grid1(ne);
    loss_kernel<<<gl, 256, 0, n->stream>>>((const float *)x_dev, (const float *)t_dev, ne, 1.0f, n->d0, n->lpart);
    NCHK(hipMemcpyAsync(n->h_lpart, n->lpart, sizeof(double) * gl, hipMemcpyDeviceToHost, n->stream));
    NTRY(nsync(n));
    double sum = 0.0;
    for (unsigned b = 0; b < gl; ++b) sum += n->h_lpart[b];
    *mse = sum / (double)ne;
    return QX_OK;
}

qx_status qx_net_set_output_scale(qx_net *n, float s)
{
    if (!n || !(s > 0.0f)) return nfail(QX_EINVAL, "output scale must be positive");
    n->out_scale = s;
    return QX_OK;
}
float qx_net_output_scale(const qx_net *n) { return n ? n->out_scale : 0.0f; }

qx_status qx_net_set_input_scale(qx_net *n, float s)
{
    if (!n || !(s >= 0.0f)) return nfail(QX_EINVAL, "input scale must be >= 0");
    n->in_scale = s;
    return QX_OK;
}
float qx_net_input_scale(const qx_net *n) { return n ? n->in_scale : 0.0f; }

static qx_status check_ctx(qx_net *n, qx_ctx *ctx, float **psi, float **V)
{
    void *p = nullptr, *v = nullptr;
    if (!n || !ctx) return nfail(QX_EINVAL, "null argument");
    if (qx_sites(ctx) != n->P || qx_precision_of(ctx) != QX_FP32 || qx_dim(ctx) != 2)
        return nfail(QX_EINVAL, "the corrector needs a 2D FP32 lattice of the network's size");
    if (qx_device_state(ctx, &p, &v) != QX_OK) return nfail(QX_EINVAL, qx_last_error());
    *psi = (float *)p;
    *V = (float *)v;
    return QX_OK;
}

static void lattice_input(qx_net *n, const float *psi, const float *V, hipStream_t s)
{
    input_kernel<<<grid1(n->P), 256, 0, s>>>(psi, V, n->in_scale, n->P, n->xin);
}

// the corrected state y scaled to sum |psi|^2 = target, written to psi
static qx_status renormalize_into(qx_net *n, float *psi, double target, hipStream_t s)
{
    const unsigned nb = std::min<unsigned>(grid1(n->P), 1024);
    norm2_partial<<<nb, 256, 0, s>>>(n->y, n->P, n->lpart);
    renorm_copy<<<grid1(2 * n->P), 256, 0, s>>>(n->y, n->lpart, (int)nb, target, 2 * n->P, psi);
    NCHK(hipGetLastError());
    return QX_OK;
}

// sum |psi|^2 of a device psi array (FP64, partials summed on the host)
static qx_status norm2_host(qx_net *n, const float *psi, hipStream_t s, double *out)
{
    const unsigned nb = std::min<unsigned>(grid1(n->P), 1024);
    norm2_partial<<<nb, 256, 0, s>>>(psi, n->P, n->lpart);
    NCHK(hipMemcpyAsync(n->h_lpart, n->lpart, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    NCHK(hipStreamSynchronize(s));
    double t = 0.0;
    for (unsigned b = 0; b < nb; ++b) t += n->h_lpart[b];
    *out = t;
    return QX_OK;
}

// psi <- psi + net(psi, V) on a 2D FP32 lattice, scaled back to its mass
// before the correction; V is not updated
qx_status qx_net_correct(qx_net *n, qx_ctx *ctx)
{
    float *psi, *V;
    NTRY(check_ctx(n, ctx, &psi, &V));
    NCHK(hipSetDevice(n->device));
    NCHK(hipStreamSynchronize((hipStream_t)qx_stream(ctx)));
    double target = 0.0;
    NTRY(norm2_host(n, psi, n->stream, &target));
    lattice_input(n, psi, V, n->stream);
    NTRY(forward(n, n->xin, n->y, false));
    NTRY(renormalize_into(n, psi, target, n->stream));
    return nsync(n);
}

qx_status qx_net_set_energy_tol(qx_net *n, double tol)
{
    if (!n || !(tol >= 0.0)) return nfail(QX_EINVAL, "energy tolerance must be >= 0");
    n->energy_tol = tol;
    return QX_OK;
}

// nsteps corrected coarse steps of length k*dt, each: leapfrog(k dt, 1),
// psi += net(psi), psi scaled back to the mass it had when the call began
// (the steps are unitary, so that is its mass), potential update; all
// stream-ordered on the lattice's stream.  Every `check_every` steps the
// relative mass change over the chunk, and with an energy tolerance set
// (qx_net_set_energy_tol) the relative change of the energy, are read back;
// beyond either tolerance the chunk is redone as k*chunk fine steps from
// the saved state (*fallbacks counts redone chunks).
qx_status qx_net_coarse_steps(qx_net *n, qx_ctx *ctx, double dt, int k, int nsteps, int check_every,
                              double mass_tol, int *fallbacks)
{
    float *psi, *V;
    NTRY(check_ctx(n, ctx, &psi, &V));
    NCHK(hipSetDevice(n->device));
    if (k < 1 || nsteps < 0 || check_every < 1) return nfail(QX_EINVAL, "need k >= 1, nsteps >= 0, check_every >= 1");
    if (!n->packed) NTRY(pack(n));
    const size_t bp = sizeof(float) * n->P * 2, bv = sizeof(float) * n->P;
    if (!n->save_psi && hipMalloc(&n->save_psi, bp) != hipSuccess) return nfail(QX_ENOMEM, "save buffer");
    if (!n->save_V && hipMalloc(&n->save_V, bv) != hipSuccess) return nfail(QX_ENOMEM, "save buffer");
    if (!n->save_tick && hipMalloc((void **)&n->save_tick, 256) != hipSuccess) return nfail(QX_ENOMEM, "save buffer");
    // the drift counter (lines of 512 points and more) picks each drift's
    // twiddle table: a redone chunk starts from the saved one, so that it
    // equals plain fine stepping
    unsigned *tick = qx::drift_tick(ctx);
    hipStream_t s = (hipStream_t)qx_stream(ctx);
    int fb = 0;
    double m0 = 0.0, target = 0.0, e0 = 0.0;
    const bool egate = n->energy_tol > 0.0;
    auto energy = [&](double *e) -> qx_status {
        double m, kin, pot;
        if (qx_energy(ctx, &m, &kin, &pot) != QX_OK) return nfail(QX_EDEVICE, qx_last_error());
        *e = kin + pot;
        return QX_OK;
    };
    if (qx_density_stats(ctx, nullptr, &m0) != QX_OK) return nfail(QX_EDEVICE, qx_last_error());
    NTRY(norm2_host(n, psi, s, &target));
    if (egate) NTRY(energy(&e0));
    for (int done = 0; done < nsteps;) {
        const int chunk = std::min(check_every, nsteps - done);
        NCHK(hipMemcpyAsync(n->save_psi, psi, bp, hipMemcpyDeviceToDevice, s));
        NCHK(hipMemcpyAsync(n->save_V, V, bv, hipMemcpyDeviceToDevice, s));
        if (tick) NCHK(hipMemcpyAsync(n->save_tick, tick, sizeof(unsigned), hipMemcpyDeviceToDevice, s));
        for (int i = 0; i < chunk; ++i) {
            if (qx::leapfrog_enqueue(ctx, k * dt, 1, 1, 1) != QX_OK) return nfail(QX_EDEVICE, qx_last_error());
            lattice_input(n, psi, V, s);
            NTRY(forward(n, n->xin, n->y, false, s));
            NTRY(renormalize_into(n, psi, target, s));
            if (qx::update_potential_enqueue(ctx) != QX_OK) return nfail(QX_EDEVICE, qx_last_error());
        }
        double m1 = 0.0, e1 = 0.0;
        if (qx_density_stats(ctx, nullptr, &m1) != QX_OK) return nfail(QX_EDEVICE, qx_last_error());
        if (egate) NTRY(energy(&e1));
        const bool bad = !(std::fabs(m1 - m0) <= mass_tol * m0) ||
                         (egate && !(std::fabs(e1 - e0) <= n->energy_tol * std::fabs(e0)));
        if (bad) {
            ++fb;
            NCHK(hipMemcpyAsync(psi, n->save_psi, bp, hipMemcpyDeviceToDevice, s));
            NCHK(hipMemcpyAsync(V, n->save_V, bv, hipMemcpyDeviceToDevice, s));
            if (tick) NCHK(hipMemcpyAsync(tick, n->save_tick, sizeof(unsigned), hipMemcpyDeviceToDevice, s));
            if (qx_leapfrog(ctx, dt, k * chunk, 1, 1) != QX_OK) return nfail(QX_EDEVICE, qx_last_error());
            if (qx_density_stats(ctx, nullptr, &m1) != QX_OK) return nfail(QX_EDEVICE, qx_last_error());
            if (egate) NTRY(energy(&e1));
        }
        m0 = m1;
        e0 = e1;
        done += chunk;
    }
    if (fallbacks) *fallbacks = fb;
    return qx_synchronize(ctx) == QX_OK ? QX_OK : nfail(QX_EDEVICE, qx_last_error());
}

}  // extern "C"
