/*
 * qx_oracle.h -- FP64 CPU restatement of qaxions' kick-drift-kick
 * Schroedinger-Poisson step.  TEST INFRASTRUCTURE ONLY: the oracle is what
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg compare
 * the HIP library against.  No product code links it.
 *
 * Conventions (all follow the reference):
 *   - lattices are N^dim, dim in {2,3}, C row-major, last index fastest;
 *   - psi is S = N^dim complex128 values, interleaved (re, im);
 *   - V is S doubles; Vhat is the half spectrum H = N^(dim-1)*(N/2+1)
 *     complex128 values laid out [i0][i1][k_last] (src/fft/fft.h:63-64);
 *   - transforms are unnormalised with FFTW signs: forward e^{-2 pi i jk/N},
 *     backward e^{+2 pi i jk/N} (src/fft/fft.h:93-104);
 *   - wavenumbers k = (2 pi / L) n with n folded to (-N/2, N/2], Nyquist at
 *     +N/2 (src/field.cxx:145-169).
 *
 * Parity status of each entry point is recorded beside it ("pinned by ...").
 */
#ifndef QX_ORACLE_H
#define QX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- transforms (src/fft/fft.h:122-140) --------------------------------
 * pinned by: brute-force O(N^2) DFT, numpy.fft, impulse/plane-wave known
 * answers, Parseval, round trip = S * identity (tests/test_oracle_fft.py). */
void qxo_fft_c2c(int dim, int N, double *data, int sign);   /* in place */
void qxo_fft_r2c(int dim, int N, const double *in, double *out_half);
void qxo_fft_c2r(int dim, int N, const double *in_half, double *out);
void qxo_fft_line(int n, double *line, int sign);           /* one 1-D line */

/* ---- split-step operators (src/field.cxx) ------------------------------ */
/* kick: psi <- psi * exp(-i dt V)  (src/field.cxx:93-116) */
void qxo_kick(int dim, int N, double *psi, const double *V, double dt);
/* drift: psi <- IFFT(exp(-i dt k^2/2)/S * FFT(psi)) (src/field.cxx:119-198) */
void qxo_drift(int dim, int N, double L, double *psi, double dt);
/* Poisson: Vhat <- Vhat * (-1/(k^2 S)); k0_mode 1 = reference (k=0 mode
 * multiplied by -1/S, src/field.cxx:253,282), 0 = k=0 mode zeroed. */
void qxo_poisson(int dim, int N, double L, double *Vhat, int k0_mode);
/* updatePotential (src/field.cxx:294-346): V <- pref(|psi|^2 - rho_bar),
 * r2c, Poisson, c2r, returns max|V|.  rho_bar_mode 1 uses <|psi|^2>. */
double qxo_update_potential(int dim, int N, double L, const double *psi,
                            double *V, double *Vhat, double pref,
                            double rho_bar, int k0_mode, int rho_bar_mode);
/* computeEnergy (src/field.cxx:200-227): V <- |psi|^2, returns max/mean */
double qxo_compute_energy(int dim, int N, const double *psi, double *V);
/* density statistics without touching V: rho_max/<rho> and sum |psi|^2 */
void qxo_density_stats(int dim, int N, const double *psi, double *rhomax_over_mean,
                       double *sum_rho);

/* ---- integrals (additions with no reference counterpart) --------------
 * mass = dV sum|psi|^2;  kinetic = 1/2 dV/S sum_k k^2 |psihat_k|^2;
 * potential = 1/2 dV sum V |psi|^2, dV = (L/N)^dim. */
void qxo_mass_energy(int dim, int N, double L, const double *psi, const double *V,
                     double *mass, double *kinetic, double *potential);

/* ---- power spectrum (src/spectrum/spectrum.cxx:32-164) ----------------
 * Spectrum of the real field `field` (the reference uses V as it stands).
 * Shell b = round(|n|), shells >= N/2+1 dropped, weight 1 for k_last in
 * {0, N/2} else 2, P[b] += w |Vhat|^2 / S^2, then P[b] /= count[b].
 * Output bins: bin j gathers shells [j*spb, (j+1)*spb); shells beyond
 * nbins*spb are dropped.  Reference mode is nbins = N/2+1, spb = 1.
 * kmodes[j] = j*spb*2pi/L (src/spectrum/spectrum.cxx:159-164). */
void qxo_spectrum(int dim, int N, double L, const double *field, int nbins, int spb,
                  double *Pk, int64_t *count, double *kmodes);

/* ---- step compositions ----------------------------------------------- */
typedef struct {
    double pref;      /* norm_ * a  (src/field.cxx:298)                  */
    double rho_bar;   /* rho_mean_ (src/field.cxx:78)                    */
    int k0_mode;      /* 1 reference, 0 zeroed                           */
    int rho_bar_mode; /* 0 fixed rho_bar, 1 actual mean                  */
} qxo_physics;

/* One merged-kick leapfrog segment (src/main.cxx:65-85): requires V current.
 * Opens with K(ds/2) if first_half else K(ds); then nsteps x [D(ds), P]
 * with K(ds) between steps; closes with K(ds/2) if last_half. */
void qxo_leapfrog(int dim, int N, double L, const qxo_physics *ph, double *psi,
                  double *V, double *Vhat, double ds, int nsteps, int first_half,
                  int last_half);
/* MRE cosmology (src/field.cxx:82-89, 349-361) unrolled into per-step
 * constants for qxo_leapfrog_schedule: drift_dt[n], kick_dt[n+1], pref[n] */
void qxo_mre_schedule(double ai, double dt, int nsteps, int first_half, int last_half,
                      double *drift_dt, double *kick_dt, double *pref, double *a_out,
                      double *s_out);
/* merged-kick leapfrog with per-step drift dt, potential prefactor and kick */
void qxo_leapfrog_schedule(int dim, int N, double L, const qxo_physics *ph, double *psi,
                           double *V, double *Vhat, int nsteps, const double *drift_dt,
                           const double *kick_dt, const double *pref);
/* propagate(): unmerged K/2 D P K/2 (src/field.cxx:388-402) */
void qxo_propagate(int dim, int N, double L, const qxo_physics *ph, double *psi,
                   double *V, double *Vhat, double ds);

/* measurement schedule (src/utils.cxx:8-28); returns count written */
int qxo_meas_list(int64_t nsteps, int64_t nmeas, int64_t *out);

/* ---- initial conditions ------------------------------------------------ */
/* SolitonsIC::apply (src/ic/solitons.cxx:20-110); sol = rows of
 * (amp, sigma, x, y, z) in box units. */
void qxo_ic_solitons(int dim, int N, double L, double bkg, const double *sol, int nsol,
                     double *psi);
/* counter-based normal deviates shared bit-for-bit with the HIP library */
uint64_t qxo_rng_u64(uint64_t seed, uint64_t stream, uint64_t idx);
double qxo_rng_uniform(uint64_t seed, uint64_t stream, uint64_t idx);
double qxo_rng_normal(uint64_t seed, uint64_t stream, uint64_t idx);
/* Gaussian random field IC: delta with P(k) ~ k^-2 for 0 < |n| <= N/4,
 * psi = sqrt(max(0, 1 + amp delta/sigma)) exp(i theta), theta = low-passed
 * U[0, 2 pi) noise at the same cutoff. */
void qxo_ic_grf(int dim, int N, double L, uint64_t seed, double amp, double *psi);
/* Soliton blobs: psi = sqrt(bkg) + sum_j sqrt(rc_j) exp(-r_j^2/(4 w_j^2)) e^{i phi_j}
 * with minimum-image r_j; blob table rows (rho_c, r_c, x, y, z, phi). */
void qxo_ic_blob_table(int nblobs, uint64_t seed, double L, double *table);
void qxo_ic_blobs(int dim, int N, double L, double bkg, const double *table, int nblobs,
                  double *psi);

/* threads used by the OpenMP loops (0 = runtime default) */
void qxo_set_threads(int nthr);
int qxo_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
