/*
 * qx_oracle.c -- FP64 CPU restatement of the qaxions KDK Schroedinger-Poisson
 * hot path.  TEST INFRASTRUCTURE ONLY (results are compared against it): tests/, smoke() and the
 * bench's cpu_baseline leg may call it, product code never does.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * The FFT is this file's own radix-2 (power-of-two N) or direct O(N^2) DFT
 * (any other N); it stands in for FFTW 3, which the reference links
 * (src/fft/fft.h:93-139) but which is absent from this environment.
 */
#include "qx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TWO_PI 6.283185307179586476925286766559

static int g_threads = 0;

void qxo_set_threads(int nthr) { g_threads = nthr; }
int qxo_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#define OMP_THREADS num_threads(qxo_get_threads())

static size_t ipow(int N, int d) { size_t s = 1; while (d-- > 0) s *= (size_t)N; return s; }
static int fold(int i, int N) { return (i <= N / 2) ? i : i - N; }  /* src/field.cxx:151-153 */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* ------------------------------------------------------------------------
 * 1-D transforms
 * ---------------------------------------------------------------------- */
typedef struct { int n; double *w; } twiddles;   /* w[2m] = cos, w[2m+1] = -sin of 2 pi m/n */
#define TW_CACHE 16
static twiddles g_tw[TW_CACHE];
static int g_ntw = 0;

static const double *get_twiddles(int n)
{
    for (int i = 0; i < g_ntw; ++i) if (g_tw[i].n == n) return g_tw[i].w;
    double *w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int m = 0; m < n; ++m) {
        /* exact multiples of pi/2 first, so symmetric entries are exact */
        double a = TWO_PI * (double)m / (double)n;
        w[2 * m] = cos(a);
        w[2 * m + 1] = -sin(a);
        if (4 * m == n) { w[2 * m] = 0.0; w[2 * m + 1] = -1.0; }
        if (2 * m == n) { w[2 * m] = -1.0; w[2 * m + 1] = 0.0; }
        if (4 * m == 3 * n) { w[2 * m] = 0.0; w[2 * m + 1] = 1.0; }
    }
    if (g_ntw < TW_CACHE) { g_tw[g_ntw].n = n; g_tw[g_ntw].w = w; ++g_ntw; }
    else { free(g_tw[0].w); g_tw[0].n = n; g_tw[0].w = w; }
    return w;
}

/* in-place unnormalised DFT of one contiguous interleaved line; sign -1 is
 * FFTW_FORWARD, +1 FFTW_BACKWARD.  tw must come from get_twiddles(n). */
static void line_fft(int n, double *a, int sign, const double *tw, double *scratch)
{
    if (n == 1) return;
    if (is_pow2(n)) {
        /* bit reversal */
        for (int i = 1, j = 0; i < n; ++i) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double t0 = a[2 * i], t1 = a[2 * i + 1];
                a[2 * i] = a[2 * j]; a[2 * i + 1] = a[2 * j + 1];
                a[2 * j] = t0; a[2 * j + 1] = t1;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int half = len >> 1, step = n / len;
            for (int i = 0; i < n; i += len) {
                for (int j = 0; j < half; ++j) {
                    double wr = tw[2 * (j * step)], wi = tw[2 * (j * step) + 1];
                    if (sign > 0) wi = -wi;
                    double *u = a + 2 * (i + j), *v = a + 2 * (i + j + half);
                    double vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;
                    v[0] = u[0] - vr; v[1] = u[1] - vi;
                    u[0] += vr; u[1] += vi;
                }
            }
        }
        return;
    }
    /* direct DFT for any other length */
    for (int k = 0; k < n; ++k) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < n; ++j) {
            int m = (int)(((long long)j * k) % n);
            double wr = tw[2 * m], wi = tw[2 * m + 1];
            if (sign > 0) wi = -wi;
            sr += a[2 * j] * wr - a[2 * j + 1] * wi;
            si += a[2 * j] * wi + a[2 * j + 1] * wr;
        }
        scratch[2 * k] = sr; scratch[2 * k + 1] = si;
    }
    memcpy(a, scratch, sizeof(double) * 2 * (size_t)n);
}

void qxo_fft_line(int n, double *line, int sign)
{
    const double *tw = get_twiddles(n);
    double *scratch = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    line_fft(n, line, sign, tw, scratch);
    free(scratch);
}

/* Transform every line along `axis` of an array of complex values with
 * shape shp[0..nd-1] (row-major).  Columns are gathered in blocks of up to
 * 16 so strided axes read whole cache lines. */
static void fft_axis(double *data, int nd, const int *shp, int axis, int sign)
{
    size_t inner = 1, outer = 1;
    for (int d = axis + 1; d < nd; ++d) inner *= (size_t)shp[d];
    for (int d = 0; d < axis; ++d) outer *= (size_t)shp[d];
    const int n = shp[axis];
    const double *tw = get_twiddles(n);
    const size_t B = inner >= 16 ? 16 : inner;
    const size_t nblk = (inner + B - 1) / B;
    const long long total = (long long)(outer * nblk);

#pragma omp parallel OMP_THREADS
    {
        double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)n * B);
        double *scr = (double *)malloc(sizeof(double) * 2 * (size_t)n);
#pragma omp for schedule(static)
        for (long long t = 0; t < total; ++t) {
            size_t o = (size_t)t / nblk, c0 = ((size_t)t % nblk) * B;
            size_t nb = (c0 + B <= inner) ? B : inner - c0;
            double *base = data + 2 * (o * (size_t)n * inner + c0);
            for (int j = 0; j < n; ++j)
                for (size_t c = 0; c < nb; ++c) {
                    buf[2 * (c * n + j)] = base[2 * ((size_t)j * inner + c)];
                    buf[2 * (c * n + j) + 1] = base[2 * ((size_t)j * inner + c) + 1];
                }
            for (size_t c = 0; c < nb; ++c) line_fft(n, buf + 2 * c * n, sign, tw, scr);
            for (int j = 0; j < n; ++j)
                for (size_t c = 0; c < nb; ++c) {
                    base[2 * ((size_t)j * inner + c)] = buf[2 * (c * n + j)];
                    base[2 * ((size_t)j * inner + c) + 1] = buf[2 * (c * n + j) + 1];
                }
        }
        free(buf); free(scr);
    }
}

/* fftw_plan_dft_{2d,3d} executed in place (src/fft/fft.h:93-94,100-101,122-130) */
void qxo_fft_c2c(int dim, int N, double *data, int sign)
{
    int shp[3] = {N, N, N};
    get_twiddles(N);  /* build outside the parallel region */
    for (int ax = dim - 1; ax >= 0; --ax) fft_axis(data, dim, shp, ax, sign);
}

/* fftw_plan_dft_r2c_{2d,3d} (src/fft/fft.h:95,102,132-135): half spectrum on
 * the last axis, H = N^(dim-1) (N/2+1).  Computed through a full complex
 * transform of the real input. */
void qxo_fft_r2c(int dim, int N, const double *in, double *out)
{
    const size_t S = ipow(N, dim), rows = S / (size_t)N, h = (size_t)N / 2 + 1;
    double *full = (double *)malloc(sizeof(double) * 2 * S);
#pragma omp parallel for OMP_THREADS
    for (long long i = 0; i < (long long)S; ++i) { full[2 * i] = in[i]; full[2 * i + 1] = 0.0; }
    qxo_fft_c2c(dim, N, full, -1);
#pragma omp parallel for OMP_THREADS
    for (long long r = 0; r < (long long)rows; ++r)
        memcpy(out + 2 * (size_t)r * h, full + 2 * (size_t)r * N, sizeof(double) * 2 * h);
    free(full);
}

/* fftw_plan_dft_c2r_{2d,3d} (src/fft/fft.h:96,103,137-140): the input is
 * extended to the full Hermitian spectrum X[-k] = conj X[k]; the output is
 * the real part of the unnormalised backward transform.  The input is not
 * modified (FFTW may destroy it; nothing in the reference reads it after). */
void qxo_fft_c2r(int dim, int N, const double *in, double *out)
{
    const size_t S = ipow(N, dim), h = (size_t)N / 2 + 1;
    double *full = (double *)malloc(sizeof(double) * 2 * S);
    const size_t rows = S / (size_t)N;
#pragma omp parallel for OMP_THREADS
    for (long long r = 0; r < (long long)rows; ++r) {
        int i0, i1;
        if (dim == 3) { i0 = (int)(r / N); i1 = (int)(r % N); }
        else { i0 = 0; i1 = (int)r; }
        for (int k = 0; k < N; ++k) {
            double re, im;
            if ((size_t)k < h) {
                re = in[2 * ((size_t)r * h + k)];
                im = in[2 * ((size_t)r * h + k) + 1];
            } else {
                /* mirror row: (-i0, -i1) mod N, column N - k */
                size_t mr;
                if (dim == 3) mr = (size_t)((N - i0) % N) * N + (size_t)((N - i1) % N);
                else mr = (size_t)((N - i1) % N);
                re = in[2 * (mr * h + (size_t)(N - k))];
                im = -in[2 * (mr * h + (size_t)(N - k)) + 1];
            }
            full[2 * ((size_t)r * N + k)] = re;
            full[2 * ((size_t)r * N + k) + 1] = im;
        }
    }
    qxo_fft_c2c(dim, N, full, +1);
#pragma omp parallel for OMP_THREADS
    for (long long i = 0; i < (long long)S; ++i) out[i] = full[2 * i];
    free(full);
}

/* ------------------------------------------------------------------------
 * split-step operators
 * ---------------------------------------------------------------------- */

/* Field::kick (src/field.cxx:93-116) */
void qxo_kick(int dim, int N, double *psi, const double *V, double dt)
{
    const size_t S = ipow(N, dim);
    const double fac = -dt;
#pragma omp parallel for OMP_THREADS
    for (long long i = 0; i < (long long)S; ++i) {
        double ph = fac * V[i];
        double c = cos(ph), s = sin(ph);
        double re = psi[2 * i], im = psi[2 * i + 1];
        psi[2 * i] = re * c - im * s;
        psi[2 * i + 1] = re * s + im * c;
    }
}

/* Field::drift (src/field.cxx:119-198): forward c2c, multiply by
 * exp(-i dt k^2/2)/S with folded n, backward c2c. */
void qxo_drift(int dim, int N, double L, double *psi, double dt)
{
    const size_t S = ipow(N, dim);
    const double fac = -0.5 * dt, dk = TWO_PI / L;
    qxo_fft_c2c(dim, N, psi, -1);
#pragma omp parallel for OMP_THREADS
    for (long long i = 0; i < (long long)S; ++i) {
        /* summation order as in the reference: 3D slowest axis first
         * (src/field.cxx:147-171), 2D fastest axis first (:161-171) */
        double k2;
        if (dim == 3) {
            double kz = fold((int)(i % N), N) * dk, ky = fold((int)((i / N) % N), N) * dk;
            double kx = fold((int)(i / ((long long)N * N)), N) * dk;
            k2 = kx * kx + ky * ky + kz * kz;
        } else {
            double kx = fold((int)(i % N), N) * dk, ky = fold((int)(i / N), N) * dk;
            k2 = kx * kx + ky * ky;
        }
        double ph = fac * k2;
        double cs = cos(ph), sn = sin(ph);
        double re = psi[2 * i], im = psi[2 * i + 1];
        psi[2 * i] = (re * cs - im * sn) / (double)S;
        psi[2 * i + 1] = (re * sn + im * cs) / (double)S;
    }
    qxo_fft_c2c(dim, N, psi, +1);
}

/* Field::Poisson (src/field.cxx:230-291): the half-axis index is the last
 * one and is not folded; k^2 == 0 becomes 1 (reference) or the mode is
 * zeroed (k0_mode 0). */
void qxo_poisson(int dim, int N, double L, double *Vhat, int k0_mode)
{
    const int h1 = N / 2 + 1;
    const double twopi = TWO_PI / L;
    const double vol = (double)ipow(N, dim);
    const size_t H = ipow(N, dim - 1) * (size_t)h1;
#pragma omp parallel for OMP_THREADS
    for (long long i = 0; i < (long long)H; ++i) {
        int nx = (int)(i % h1);
        long long r = i / h1;
        int ny = fold((int)(r % N), N);
        int nz = (dim == 3) ? fold((int)(r / N), N) : 0;
        double kx = twopi * nx, ky = twopi * ny, kz = twopi * nz;
        double k2 = kx * kx + ky * ky + kz * kz;
        double f;
        if (k2 == 0.0) f = k0_mode ? -1.0 / 1.0 / vol : 0.0;
        else f = -1.0 / k2 / vol;
        Vhat[2 * i] *= f;
        Vhat[2 * i + 1] *= f;
    }
}

/* Field::updatePotential (src/field.cxx:294-346) */
double qxo_update_potential(int dim, int N, double L, const double *psi, double *V,
                            double *Vhat, double pref, double rho_bar, int k0_mode,
                            int rho_bar_mode)
{
    const size_t S = ipow(N, dim);
    if (rho_bar_mode) {
        double s = 0.0;
        for (size_t i = 0; i < S; ++i) s += psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        rho_bar = s / (double)S;
    }
#pragma omp parallel for OMP_THREADS
    for (long long i = 0; i < (long long)S; ++i) {
        double rho = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        V[i] = pref * (rho - rho_bar);
    }
    qxo_fft_r2c(dim, N, V, Vhat);
    qxo_poisson(dim, N, L, Vhat, k0_mode);
    qxo_fft_c2r(dim, N, Vhat, V);
    double vmax = 0.0;
    for (size_t i = 0; i < S; ++i) if (fabs(V[i]) > vmax) vmax = fabs(V[i]);
    return vmax;
}

/* Field::computeEnergy (src/field.cxx:200-227): overwrites V with |psi|^2 */
double qxo_compute_energy(int dim, int N, const double *psi, double *V)
{
    const size_t S = ipow(N, dim);
    double mx = 0.0, sum = 0.0;
    for (size_t i = 0; i < S; ++i) {
        V[i] = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        double a = fabs(V[i]);
        if (a > mx) mx = a;
        sum += a;
    }
    return mx / (sum / (double)S);
}

void qxo_density_stats(int dim, int N, const double *psi, double *rmax_over_mean, double *sum_rho)
{
    const size_t S = ipow(N, dim);
    double mx = 0.0, sum = 0.0;
    for (size_t i = 0; i < S; ++i) {
        double r = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        if (r > mx) mx = r;
        sum += r;
    }
    if (rmax_over_mean) *rmax_over_mean = mx / (sum / (double)S);
    if (sum_rho) *sum_rho = sum;
}

void qxo_mass_energy(int dim, int N, double L, const double *psi, const double *V,
                     double *mass, double *kinetic, double *potential)
{
    const size_t S = ipow(N, dim);
    const double dV = pow(L / N, dim), dk = TWO_PI / L;
    double m = 0.0, p = 0.0;
    for (size_t i = 0; i < S; ++i) {
        double r = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        m += r;
        if (V) p += V[i] * r;
    }
    if (mass) *mass = dV * m;
    if (potential) *potential = 0.5 * dV * p;
    if (kinetic) {
        double *h = (double *)malloc(sizeof(double) * 2 * S);
        memcpy(h, psi, sizeof(double) * 2 * S);
        qxo_fft_c2c(dim, N, h, -1);
        double k = 0.0;
        for (size_t i = 0; i < S; ++i) {
            long long q = (long long)i;
            int a = fold((int)(q % N), N), b = fold((int)((q / N) % N), N);
            int c = dim == 3 ? fold((int)(q / ((long long)N * N)), N) : 0;
            double k2 = dk * dk * (double)(a * a + b * b + c * c);
            k += k2 * (h[2 * i] * h[2 * i] + h[2 * i + 1] * h[2 * i + 1]);
        }
        *kinetic = 0.5 * dV / (double)S * k;
        free(h);
    }
}

/* Spectrum::compute/bin2D/bin3D/buildKAxis (src/spectrum/spectrum.cxx:32-164) */
void qxo_spectrum(int dim, int N, double L, const double *field, int nbins, int spb,
                  double *Pk, int64_t *count, double *kmodes)
{
    const size_t S = ipow(N, dim);
    const int hN = N / 2, h1 = N / 2 + 1, nshell = N / 2 + 1;
    const size_t H = ipow(N, dim - 1) * (size_t)h1;
    double *hat = (double *)malloc(sizeof(double) * 2 * H);
    qxo_fft_r2c(dim, N, field, hat);
    const double norm = 1.0 / ((double)S * (double)S);
    for (int b = 0; b < nbins; ++b) { Pk[b] = 0.0; count[b] = 0; }
    for (size_t i = 0; i < H; ++i) {
        int nx = (int)(i % h1);
        size_t r = i / h1;
        int ny = fold((int)(r % N), N);
        int nz = dim == 3 ? fold((int)(r / N), N) : 0;
        int shell = (int)lround(sqrt((double)(nx * nx + ny * ny + nz * nz)));
        if (shell >= nshell) continue;
        int b = shell / spb;
        if (b >= nbins) continue;
        double mult = (nx == 0 || nx == hN) ? 1.0 : 2.0;
        double re = hat[2 * i], im = hat[2 * i + 1];
        Pk[b] += mult * norm * (re * re + im * im);
        count[b] += (int64_t)mult;
    }
    for (int b = 0; b < nbins; ++b) {
        if (count[b] > 0) Pk[b] /= (double)count[b];
        kmodes[b] = ((double)b * spb + 0.5 * (spb - 1)) * (TWO_PI / L);
    }
    free(hat);
}

/* ------------------------------------------------------------------------
 * step compositions
 * ---------------------------------------------------------------------- */
static void update_pot(int dim, int N, double L, const qxo_physics *ph, double *psi, double *V,
                       double *Vhat)
{
    qxo_update_potential(dim, N, L, psi, V, Vhat, ph->pref, ph->rho_bar, ph->k0_mode,
                         ph->rho_bar_mode);
}

/* src/main.cxx:65-85 with Field::half_kick/full_kick/drift_update
 * (src/field.cxx:363-386) */
void qxo_leapfrog(int dim, int N, double L, const qxo_physics *ph, double *psi, double *V,
                  double *Vhat, double ds, int nsteps, int first_half, int last_half)
{
    qxo_kick(dim, N, psi, V, first_half ? 0.5 * ds : ds);
    for (int i = 0; i < nsteps; ++i) {
        qxo_drift(dim, N, L, psi, ds);
        update_pot(dim, N, L, ph, psi, V, Vhat);
        if (i != nsteps - 1) qxo_kick(dim, N, psi, V, ds);
    }
    if (last_half) qxo_kick(dim, N, psi, V, 0.5 * ds);
}

/* Field::setCosmo MRE branch (src/field.cxx:82-89) and Field::updateTime
 * (src/field.cxx:349-361) unrolled over the merged-kick loop of
 * src/main.cxx:65-85: drift_dt[j] = ds before step j's updateTime, pref[j] =
 * norm_ a after it (norm_ = 1 in MRE, src/field.cxx:88, pref at :298),
 * kick_dt[0] the opening kick, kick_dt[j] (0 < j < n) the merged kick
 * between steps j-1 and j, kick_dt[n] the closing kick (0 when absent).
 * a_out/s_out: a and s after the n steps. */
void qxo_mre_schedule(double ai, double dt, int nsteps, int first_half, int last_half,
                      double *drift_dt, double *kick_dt, double *pref, double *a_out,
                      double *s_out)
{
    double a = ai, ds = dt / ai, s = 0.0;
    kick_dt[0] = first_half ? 0.5 * ds : ds;
    for (int j = 0; j < nsteps; ++j) {
        drift_dt[j] = ds;
        ds = 0.00025 / a;
        a = a * (1 + 1.27 * 1.27 * sqrt(1 + a) * ds);
        s += ds;
        pref[j] = a;
        kick_dt[j + 1] = (j != nsteps - 1) ? ds : (last_half ? 0.5 * ds : 0.0);
    }
    *a_out = a;
    *s_out = s;
}

/* the merged-kick loop of src/main.cxx:65-85 with per-step constants (the
 * MRE run, where Field::updateTime changes ds and a every step): requires V
 * current; K(kick_dt[0]); then for each step j: D(drift_dt[j]), P with
 * pref[j], and K(kick_dt[j+1]) (skipped when 0) */
void qxo_leapfrog_schedule(int dim, int N, double L, const qxo_physics *ph, double *psi,
                           double *V, double *Vhat, int nsteps, const double *drift_dt,
                           const double *kick_dt, const double *pref)
{
    qxo_kick(dim, N, psi, V, kick_dt[0]);
    for (int j = 0; j < nsteps; ++j) {
        qxo_physics pj = *ph;
        pj.pref = pref[j];
        qxo_drift(dim, N, L, psi, drift_dt[j]);
        update_pot(dim, N, L, &pj, psi, V, Vhat);
        if (kick_dt[j + 1] != 0.0) qxo_kick(dim, N, psi, V, kick_dt[j + 1]);
    }
}

/* Field::propagate (src/field.cxx:388-402) */
void qxo_propagate(int dim, int N, double L, const qxo_physics *ph, double *psi, double *V,
                   double *Vhat, double ds)
{
    qxo_kick(dim, N, psi, V, 0.5 * ds);
    qxo_drift(dim, N, L, psi, ds);
    update_pot(dim, N, L, ph, psi, V, Vhat);
    qxo_kick(dim, N, psi, V, 0.5 * ds);
}

/* generateMeasList (src/utils.cxx:8-28) */
int qxo_meas_list(int64_t nsteps, int64_t nmeas, int64_t *out)
{
    if (nmeas <= 0 || nsteps <= 0) return 0;
    if (nmeas == 1) { out[0] = nsteps - 1; return 1; }
    for (int64_t k = 0; k < nmeas; ++k) out[k] = k * (nsteps - 1) / (nmeas - 1);
    return (int)nmeas;
}

/* ------------------------------------------------------------------------
 * initial conditions
 * ---------------------------------------------------------------------- */

/* SolitonsIC::apply (src/ic/solitons.cxx:20-110): x is the fastest index,
 * centres are box fractions, no periodic images, rho normalised to mean 1. */
void qxo_ic_solitons(int dim, int N, double L, double bkg, const double *sol, int nsol,
                     double *psi)
{
    const size_t S = ipow(N, dim);
    const double dx = L / N;
    double *rho = (double *)malloc(sizeof(double) * S);
    double sum = 0.0;
    for (size_t i = 0; i < S; ++i) {
        double x, y, z = 0.0;
        if (dim == 3) {
            int iz = (int)(i / ((size_t)N * N)), iy = (int)((i / N) % N), ix = (int)(i % N);
            x = ix * dx; y = iy * dx; z = iz * dx;
        } else {
            int iy = (int)(i / N), ix = (int)(i % N);
            x = ix * dx; y = iy * dx;
        }
        double r = bkg;
        for (int s = 0; s < nsol; ++s) {
            double amp = sol[5 * s], sg = sol[5 * s + 1];
            double nrm = 1.0 / (sg * sg * sg * 2.0 * M_PI);
            double ddx = x - sol[5 * s + 2] * L, ddy = y - sol[5 * s + 3] * L;
            double ddz = dim == 3 ? (z - sol[5 * s + 4] * L) : 0.0;
            double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            r += amp * nrm * exp(-r2 / (2.0 * sg * sg));
        }
        rho[i] = r;
        sum += r;
    }
    double mean = sum / (double)S;
    for (size_t i = 0; i < S; ++i) { psi[2 * i] = sqrt(rho[i] / mean); psi[2 * i + 1] = 0.0; }
    free(rho);
}

/* SplitMix64 finaliser (Steele, Lea, Flood 2014) applied to a keyed counter */
static uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
uint64_t qxo_rng_u64(uint64_t seed, uint64_t stream, uint64_t idx)
{
    uint64_t key = mix64(seed ^ (stream * 0xD6E8FEB86659FD93ULL));
    return mix64(key + (idx + 1) * 0x9E3779B97F4A7C15ULL);
}
double qxo_rng_uniform(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return (double)(qxo_rng_u64(seed, stream, idx) >> 11) * 0x1.0p-53;
}
/* Box-Muller on two keyed draws; u1 in (0, 1] */
double qxo_rng_normal(uint64_t seed, uint64_t stream, uint64_t idx)
{
    double u1 = (double)((qxo_rng_u64(seed, stream, 2 * idx) >> 11) + 1) * 0x1.0p-53;
    double u2 = (double)(qxo_rng_u64(seed, stream, 2 * idx + 1) >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(TWO_PI * u2);
}

/* multiply a full spectrum by a radial filter; mode 0: 1/|n| for
 * 0 < |n|^2 <= (N/4)^2, mode 1: 1 for |n|^2 <= (N/4)^2 */
static void radial_filter(int dim, int N, double *hat, int mode)
{
    const size_t S = ipow(N, dim);
    const long long cut = (long long)N * N;  /* 16 |n|^2 <= N^2 */
#pragma omp parallel for OMP_THREADS
    for (long long i = 0; i < (long long)S; ++i) {
        int a = fold((int)(i % N), N), b = fold((int)((i / N) % N), N);
        int c = dim == 3 ? fold((int)(i / ((long long)N * N)), N) : 0;
        long long m = (long long)a * a + (long long)b * b + (long long)c * c;
        double f;
        if (16 * m > cut) f = 0.0;
        else if (mode == 0) f = (m == 0) ? 0.0 : 1.0 / sqrt((double)m);
        else f = 1.0;
        hat[2 * i] *= f;
        hat[2 * i + 1] *= f;
    }
}

void qxo_ic_grf(int dim, int N, double L, uint64_t seed, double amp, double *psi)
{
    (void)L;
    const size_t S = ipow(N, dim);
    double *d = (double *)malloc(sizeof(double) * 2 * S);
    double *t = (double *)malloc(sizeof(double) * 2 * S);
#pragma omp parallel for OMP_THREADS
    for (long long i = 0; i < (long long)S; ++i) {
        d[2 * i] = qxo_rng_normal(seed, 0, (uint64_t)i); d[2 * i + 1] = 0.0;
        t[2 * i] = TWO_PI * qxo_rng_uniform(seed, 1, (uint64_t)i); t[2 * i + 1] = 0.0;
    }
    qxo_fft_c2c(dim, N, d, -1);
    radial_filter(dim, N, d, 0);
    qxo_fft_c2c(dim, N, d, +1);
    qxo_fft_c2c(dim, N, t, -1);
    radial_filter(dim, N, t, 1);
    qxo_fft_c2c(dim, N, t, +1);
    double s1 = 0.0, s2 = 0.0;
    for (size_t i = 0; i < S; ++i) { double v = d[2 * i] / (double)S; s1 += v; s2 += v * v; }
    double mean = s1 / (double)S;
    double sigma = sqrt(s2 / (double)S - mean * mean);
#pragma omp parallel for OMP_THREADS
    for (long long i = 0; i < (long long)S; ++i) {
        double delta = (d[2 * i] / (double)S - mean) / sigma;
        double a2 = 1.0 + amp * delta;
        double a = sqrt(a2 > 0.0 ? a2 : 0.0);
        double th = t[2 * i] / (double)S;
        psi[2 * i] = a * cos(th);
        psi[2 * i + 1] = a * sin(th);
    }
    free(d); free(t);
}

void qxo_ic_blob_table(int nblobs, uint64_t seed, double L, double *tab)
{
    for (int j = 0; j < nblobs; ++j) {
        uint64_t st = 16 + (uint64_t)j;
        tab[6 * j + 0] = 1.0 + 9.0 * qxo_rng_uniform(seed, st, 0);
        tab[6 * j + 1] = 0.3 + 0.7 * qxo_rng_uniform(seed, st, 1);
        tab[6 * j + 2] = L * qxo_rng_uniform(seed, st, 2);
        tab[6 * j + 3] = L * qxo_rng_uniform(seed, st, 3);
        tab[6 * j + 4] = L * qxo_rng_uniform(seed, st, 4);
        tab[6 * j + 5] = TWO_PI * qxo_rng_uniform(seed, st, 5);
    }
}

void qxo_ic_blobs(int dim, int N, double L, double bkg, const double *tab, int nb, double *psi)
{
    const size_t S = ipow(N, dim);
    const double dx = L / N;
    const double sb = sqrt(bkg);
#pragma omp parallel for OMP_THREADS
    for (long long i = 0; i < (long long)S; ++i) {
        double pos[3];
        pos[0] = (double)(i % N) * dx;
        pos[1] = (double)((i / N) % N) * dx;
        pos[2] = dim == 3 ? (double)(i / ((long long)N * N)) * dx : 0.0;
        double re = sb, im = 0.0;
        for (int j = 0; j < nb; ++j) {
            const double *b = tab + 6 * j;
            double r2 = 0.0;
            for (int a = 0; a < dim; ++a) {
                double d = pos[a] - b[2 + a];
                d -= L * nearbyint(d / L);
                r2 += d * d;
            }
            double amp = sqrt(b[0]) * exp(-r2 / (4.0 * b[1] * b[1]));
            re += amp * cos(b[5]);
            im += amp * sin(b[5]);
        }
        psi[2 * i] = re;
        psi[2 * i + 1] = im;
    }
}
