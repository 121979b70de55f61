// qx_aux.hpp -- pointwise, reduction, spectrum, layout and initial-condition
// kernels that surround the FFT passes.  All are HBM-streaming kernels with a
// grid-stride loop; reductions write one FP64 partial per block, summed on
// the host in a fixed order so results are run-to-run deterministic.
#pragma once
#include "qx_kernels.hpp"

namespace qx {

constexpr int RED_THREADS = 256;

// ---- precision conversion at the host boundary ---------------------------
template <typename R>
__global__ void c128_to_c(const double2 *__restrict__ in, typename cplx<R>::T *__restrict__ out,
                          long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double2 a = in[i];
        out[i] = typename cplx<R>::T{(R)a.x, (R)a.y};
    }
}
template <typename R>
__global__ void c_to_c128(const typename cplx<R>::T *__restrict__ in, double2 *__restrict__ out,
                          long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        auto a = in[i];
        out[i] = double2{(double)a.x, (double)a.y};
    }
}
template <typename R>
__global__ void f64_to_r(const double *__restrict__ in, R *__restrict__ out, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = (R)in[i];
}
template <typename R>
__global__ void r_to_f64(const R *__restrict__ in, double *__restrict__ out, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = (double)in[i];
}

// ---- kick from the stored potential: psi <- psi exp(-i dt V) --------------
// (src/field.cxx:93-116)
template <typename R>
__global__ void kick_kernel(typename cplx<R>::T *__restrict__ psi, const R *__restrict__ V, R dt,
                            long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        R s, cm1;
        sincosm1_kick(-dt * V[i], &s, &cm1);
        psi[i] = rotate(psi[i], s, cm1);
    }
}

// ---- V <- |psi|^2 (computeEnergy, src/field.cxx:205-211) ------------------
template <typename R>
__global__ void density_kernel(const typename cplx<R>::T *__restrict__ psi, R *__restrict__ V,
                               long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        auto a = psi[i];
        V[i] = a.x * a.x + a.y * a.y;
    }
}

// ---- block reduction helpers -------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += sh[i];  // fixed order
    return s;
}
__device__ __forceinline__ double block_max(double v, double *sh)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = sh[0];
    for (int i = 1; i < nw; ++i) s = fmax(s, sh[i]);
    return s;
}

enum {
    RED_RHO = 0,   // sum |psi|^2, max |psi|^2
    RED_VRHO = 1,  // sum V |psi|^2
    RED_K2 = 2,    // sum k^2 |psihat|^2 over the full spectrum (integer m)
    RED_VABS = 3,  // sum |V|, max |V|
};

// partials[2*b] = block sum, partials[2*b+1] = block max
template <typename R, int KIND>
__global__ void __launch_bounds__(RED_THREADS)
reduce_kernel(const typename cplx<R>::T *__restrict__ psi, const R *__restrict__ V, long long n,
              int N, int dim, double *__restrict__ partials)
{
    __shared__ double sh[RED_THREADS / 64];
    double s = 0.0, mx = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        if (KIND == RED_RHO || KIND == RED_VRHO || KIND == RED_K2) {
            auto a = psi[i];
            const double r = (double)a.x * a.x + (double)a.y * a.y;
            if (KIND == RED_RHO) { s += r; mx = fmax(mx, r); }
            else if (KIND == RED_VRHO) s += (double)V[i] * r;
            else {
                const int a0 = foldi((int)(i % N), N), a1 = foldi((int)((i / N) % N), N);
                const int a2 = dim == 3 ? foldi((int)(i / ((long long)N * N)), N) : 0;
                s += (double)(a0 * a0 + a1 * a1 + a2 * a2) * r;
            }
        } else {
            const double v = fabs((double)V[i]);
            s += v;
            mx = fmax(mx, v);
        }
    }
    s = block_sum(s, sh);
    mx = block_max(mx, sh);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s;
        partials[2 * blockIdx.x + 1] = mx;
    }
}

// (sum, max) partials of reduce_kernel<., RED_RHO> summed in block order,
// as the host sums them for qx_density_stats: out[0] = max/mean of |psi|^2,
// out[1] = dV sum |psi|^2 (one thread: at most 1024 partials)
__global__ void density_stats_finalize(const double *__restrict__ partials, int blocks, double sites,
                                       double dV, double *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0, m = 0.0;
    for (int b = 0; b < blocks; ++b) {
        s += partials[2 * b];
        m = fmax(m, partials[2 * b + 1]);
    }
    out[0] = m / (s / sites);
    out[1] = dV * s;
}

// ---- power-spectrum binning (src/spectrum/spectrum.cxx:63-155) -----------
// shell = round(|n|) (shell_of), bins of `spb` consecutive shells; the power
// itself is binned inside the transform's last pass (contig_bin)

// mode counts per bin (the data-independent half of the binning): every
// half-spectrum mode adds its weight w (1 for k_last in {0, N/2}, else 2)
// to the bin of its shell
__global__ void __launch_bounds__(256)
spectrum_count_kernel(long long rows, int N, int dim, int nbins, int spb, unsigned long long *__restrict__ cnt)
{
    extern __shared__ unsigned int sc[];
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) sc[i] = 0u;
    __syncthreads();
    const int H1 = N / 2 + 1;
    const long long total = (long long)H1 * rows;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(i / rows);
        const long long row = i % rows;
        int m = k * k;
        if (dim == 3) {
            const int n0 = foldi((int)(row / N), N), n1 = foldi((int)(row % N), N);
            m += n0 * n0 + n1 * n1;
        } else {
            const int n0 = foldi((int)row, N);
            m += n0 * n0;
        }
        const int shell = shell_of(m);
        if (shell > N / 2 || shell / spb >= nbins) continue;
        atomicAdd(&sc[shell / spb], (k == 0 || k == N / 2) ? 1u : 2u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += blockDim.x)
        if (sc[i]) atomicAdd(&cnt[i], (unsigned long long)sc[i]);
}

// out[bin] = sum over the rows r < rows of part[r][bin] (the binning
// pass's block histograms): one block per bin, each thread a strided set
// of rows in order, then a fixed tree, so the sum does not depend on timing
__global__ void __launch_bounds__(256)
bin_reduce(const double *__restrict__ part, int rows, int nbins, double *__restrict__ out)
{
    __shared__ double sh[256];
    const int bin = blockIdx.x;
    double s = 0.0;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) s += part[(long long)r * nbins + bin];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[bin] = sh[0];
}

// ---- vt[k*rows + row] <-> ref[row*H1 + k] (the reference's half layout) ------
template <typename R>
__global__ void vt_to_ref(const typename cplx<R>::T *__restrict__ vt, typename cplx<R>::T *__restrict__ ref,
                          long long rows, int H1)
{
    __shared__ typename cplx<R>::T tile[32][33];
    const long long rb = (long long)blockIdx.x * 32;
    const int kb = blockIdx.y * 32;
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        const int k = kb + j;
        const long long row = rb + threadIdx.x;
        if (k < H1 && row < rows) tile[j][threadIdx.x] = vt[(long long)k * rows + row];
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        const long long row = rb + j;
        const int k = kb + threadIdx.x;
        if (k < H1 && row < rows) ref[row * H1 + k] = tile[threadIdx.x][j];
    }
}
template <typename R>
__global__ void ref_to_vt(const typename cplx<R>::T *__restrict__ ref, typename cplx<R>::T *__restrict__ vt,
                          long long rows, int H1)
{
    __shared__ typename cplx<R>::T tile[32][33];
    const long long rb = (long long)blockIdx.x * 32;
    const int kb = blockIdx.y * 32;
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        const long long row = rb + j;
        const int k = kb + threadIdx.x;
        if (k < H1 && row < rows) tile[threadIdx.x][j] = ref[row * H1 + k];
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        const int k = kb + j;
        const long long row = rb + threadIdx.x;
        if (k < H1 && row < rows) vt[(long long)k * rows + row] = tile[j][threadIdx.x];
    }
}

// ---- Poisson factor on a half spectrum in the reference layout --------------
// (src/field.cxx:230-291)
template <typename R>
__global__ void poisson_ref_kernel(typename cplx<R>::T *__restrict__ vh, long long rows, int N,
                                   int dim, double pfac, double k0fac)
{
    const int H1 = N / 2 + 1;
    const long long total = rows * H1;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int nx = (int)(i % H1);
        const long long r = i / H1;
        const int ny = foldi((int)(r % N), N);
        const int nz = dim == 3 ? foldi((int)(r / N), N) : 0;
        const long long m = (long long)nx * nx + (long long)ny * ny + (long long)nz * nz;
        const double f = (m == 0) ? k0fac : -pfac / (double)m;
        auto x = vh[i];
        vh[i] = typename cplx<R>::T{(R)(x.x * f), (R)(x.y * f)};
    }
}

// ---- counter-based random numbers (SplitMix64 finaliser on a keyed counter) --
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ unsigned long long rng_u64(unsigned long long seed, unsigned long long stream,
                                                      unsigned long long idx)
{
    const unsigned long long key = mix64(seed ^ (stream * 0xD6E8FEB86659FD93ULL));
    return mix64(key + (idx + 1) * 0x9E3779B97F4A7C15ULL);
}

// d = normal(seed, 0, i), t = 2 pi uniform(seed, 1, i)
template <typename R>
__global__ void grf_noise_kernel(typename cplx<R>::T *__restrict__ d, typename cplx<R>::T *__restrict__ t,
                                 unsigned long long seed, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const double u1 = (double)((rng_u64(seed, 0, 2 * i) >> 11) + 1) * 0x1.0p-53;
        const double u2 = (double)(rng_u64(seed, 0, 2 * i + 1) >> 11) * 0x1.0p-53;
        const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
        const double u = (double)(rng_u64(seed, 1, i) >> 11) * 0x1.0p-53;
        d[i] = typename cplx<R>::T{(R)g, (R)0};
        t[i] = typename cplx<R>::T{(R)(6.283185307179586476925286766559 * u), (R)0};
    }
}

// mode 0: 1/|n| for 0 < 16|n|^2 <= N^2; mode 1: 1 for 16|n|^2 <= N^2
template <typename R>
__global__ void radial_filter_kernel(typename cplx<R>::T *__restrict__ h, int N, int dim, int mode,
                                     long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int a = foldi((int)(i % N), N), b = foldi((int)((i / N) % N), N);
        const int c = dim == 3 ? foldi((int)(i / ((long long)N * N)), N) : 0;
        const long long m = (long long)a * a + (long long)b * b + (long long)c * c;
        R f;
        if (16 * m > (long long)N * N) f = 0;
        else if (mode == 0) f = (m == 0) ? (R)0 : (R)(1.0 / sqrt((double)m));
        else f = 1;
        auto x = h[i];
        h[i] = typename cplx<R>::T{x.x * f, x.y * f};
    }
}

// psi = sqrt(max(0, 1 + amp (d/S - mean)/sigma)) e^{i t/S}
template <typename R>
__global__ void grf_compose_kernel(typename cplx<R>::T *__restrict__ psi, const typename cplx<R>::T *__restrict__ t,
                                   double inv_s, double mean, double inv_sigma, double amp, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const double delta = ((double)psi[i].x * inv_s - mean) * inv_sigma;
        const double a2 = 1.0 + amp * delta;
        const double a = sqrt(a2 > 0.0 ? a2 : 0.0);
        const double th = (double)t[i].x * inv_s;
        double s, c;
        sincos(th, &s, &c);
        psi[i] = typename cplx<R>::T{(R)(a * c), (R)(a * s)};
    }
}

// real part statistics of d/S: partial sums of x and x^2
template <typename R>
__global__ void __launch_bounds__(RED_THREADS)
moments_kernel(const typename cplx<R>::T *__restrict__ d, double inv_s, long long n,
               double *__restrict__ partials)
{
    __shared__ double sh[RED_THREADS / 64];
    double s1 = 0.0, s2 = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const double x = (double)d[i].x * inv_s;
        s1 += x;
        s2 += x * x;
    }
    s1 = block_sum(s1, sh);
    s2 = block_sum(s2, sh);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s1;
        partials[2 * blockIdx.x + 1] = s2;
    }
}

// blobs: psi = sqrt(bkg) + sum_j sqrt(rc_j) exp(-r_j^2/(4 w_j^2)) e^{i phi_j}
template <typename R>
__global__ void blobs_kernel(typename cplx<R>::T *__restrict__ psi, const double *__restrict__ tab, int nb,
                             int N, int dim, double L, double bkg, long long n)
{
    const double dx = L / N;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double pos[3];
        pos[0] = (double)(i % N) * dx;
        pos[1] = (double)((i / N) % N) * dx;
        pos[2] = dim == 3 ? (double)(i / ((long long)N * N)) * dx : 0.0;
        double re = sqrt(bkg), im = 0.0;
        for (int j = 0; j < nb; ++j) {
            const double *b = tab + 6 * j;
            double r2 = 0.0;
            for (int a = 0; a < dim; ++a) {
                double d = pos[a] - b[2 + a];
                d -= L * rint(d / L);
                r2 += d * d;
            }
            const double amp = sqrt(b[0]) * exp(-r2 / (4.0 * b[1] * b[1]));
            double s, c;
            sincos(b[5], &s, &c);
            re += amp * c;
            im += amp * s;
        }
        psi[i] = typename cplx<R>::T{(R)re, (R)im};
    }
}

// solitons (src/ic/solitons.cxx:55-95): rho = bkg + sum amp/(2 pi s^3) e^{-r^2/2s^2}
// written to V (real) for the mean; psi = sqrt(rho/mean) afterwards
template <typename R>
__global__ void solitons_rho_kernel(R *__restrict__ rho, const double *__restrict__ sol, int ns, int N,
                                    int dim, double L, double bkg, long long n)
{
    const double dx = L / N;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const double x = (double)(i % N) * dx, y = (double)((i / N) % N) * dx;
        const double z = dim == 3 ? (double)(i / ((long long)N * N)) * dx : 0.0;
        double r = bkg;
        for (int s = 0; s < ns; ++s) {
            const double amp = sol[5 * s], sg = sol[5 * s + 1];
            const double nrm = 1.0 / (sg * sg * sg * 2.0 * 3.14159265358979323846);
            const double ddx = x - sol[5 * s + 2] * L, ddy = y - sol[5 * s + 3] * L;
            const double ddz = dim == 3 ? z - sol[5 * s + 4] * L : 0.0;
            const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            r += amp * nrm * exp(-r2 / (2.0 * sg * sg));
        }
        rho[i] = (R)r;
    }
}
template <typename R>
__global__ void solitons_psi_kernel(typename cplx<R>::T *__restrict__ psi, const R *__restrict__ rho,
                                    double inv_mean, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        psi[i] = typename cplx<R>::T{(R)sqrt((double)rho[i] * inv_mean), (R)0};
}

}  // namespace qx
