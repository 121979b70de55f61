/*
 * qx.h -- C ABI of the MI355X (gfx950) kick-drift-kick Schroedinger-Poisson
 * library (libqx.so).  Plain pointers and sizes only; no exceptions cross it.
 *
 * One qx_ctx stands for one reference `Field` (src/field.h:13-97): it owns
 * the device-resident state psi (S complex), V (S real) and the potential's
 * half spectrum, the FFT twiddles, one HIP stream and scratch.  Every call
 * returns after its work is complete on the stream, which keeps the
 * reference's synchronous semantics.  Calls on one context must be
 * serialised by the caller; distinct contexts may be used from different
 * threads.  Errors: a status code plus qx_last_error() (thread-local text).
 *
 * Host arrays are caller-owned and read/written only during the call.  At
 * this boundary psi is always complex128 interleaved (re, im), V float64,
 * half spectra complex128 in the reference layout [i0][i1][k_last] with
 * k_last in [0, N/2] (src/fft/fft.h:63-64); in QX_FP32 contexts the device
 * state is complex64/float32 and converted on the device.
 *
 * Lattices: dim 2 or 3, N a power of two in [2, 4096] (2D) or [2, 1024]
 * (3D).  Other N are rejected with QX_EINVAL.
 */
#ifndef QX_H
#define QX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qx_ctx qx_ctx;

typedef enum {
    QX_OK = 0,
    QX_EINVAL = 1,   /* bad argument (reference: std::invalid_argument)      */
    QX_ENOMEM = 2,   /* device allocation failed (reference: std::bad_alloc) */
    QX_EDEVICE = 3,  /* HIP runtime error (reference: std::runtime_error)     */
    QX_ESTATE = 4    /* call not valid in the current state                   */
} qx_status;

typedef enum { QX_FP32 = 0, QX_FP64 = 1 } qx_precision;

/* Physics constants of the step.
 *   pref     = norm_ * a (src/field.cxx:298), the Poisson source prefactor;
 *   rho_bar  = rho_mean_ (src/field.cxx:78,86), subtracted from |psi|^2;
 *   k0_mode  = 1: the k = 0 mode of Vhat is multiplied by -1/S as in the
 *              reference (src/field.cxx:253-254,282-283); 0: it is zeroed;
 *   rho_bar_mode = 1: subtract the actual mean <|psi|^2> (k = 0 zeroed). */
typedef struct {
    double pref;
    double rho_bar;
    int k0_mode;
    int rho_bar_mode;
} qx_physics;

/* ---- lifetime: Field::Field/init/~Field (src/field.cxx:13-66) ----------- */
qx_status qx_create(int dim, int N, double L, qx_precision prec, int device, qx_ctx **out);
qx_status qx_destroy(qx_ctx *ctx);
/* setCosmo STATIC branch (src/field.cxx:73-80) sets pref = 4 pi norm, rho_bar = 1 */
qx_status qx_set_physics(qx_ctx *ctx, const qx_physics *ph);
qx_status qx_get_physics(const qx_ctx *ctx, qx_physics *ph);

/* ---- state transfer: Field::psi()/V() accessors (src/field.h:36-45) ----- */
qx_status qx_upload_psi(qx_ctx *ctx, const double *psi_c128);
qx_status qx_download_psi(qx_ctx *ctx, double *psi_c128);
qx_status qx_upload_V(qx_ctx *ctx, const double *V);
qx_status qx_download_V(qx_ctx *ctx, double *V);
/* device-to-device copy of psi, V and the potential spectrum between two
 * contexts of equal shape and precision (no reference counterpart).  FP32
 * lattices with lines of 512 points and more also copy their drift counter
 * (which of the dithered twiddle tables each drift takes), so that the copy
 * steps bit for bit like the source. */
qx_status qx_copy_state(qx_ctx *dst, const qx_ctx *src);
/* device pointers of psi / V (element type set by the precision) */
qx_status qx_device_state(qx_ctx *ctx, void **psi, void **V);

/* ---- asynchronous snapshots: the host reads of IO::writeConf
 * (src/io/io.cxx:314-370, which takes V and psi from Field's host arrays)
 * without stopping the time loop.  qx_snap_take queues device copies of
 * psi and/or V on the context's stream and returns at once; their transfer
 * to pinned host memory runs on the snapshot's own stream behind the steps
 * queued after it.  qx_snap_wait blocks until that transfer is done and
 * converts to complex128 / float64 (outputs nullable); it may be called
 * from another thread than the one driving the context, and must be called
 * before the next take.  A snapshot holds S complex and/or S real of device
 * memory and as much pinned host memory.  Only take uses the context; wait
 * and destroy remain valid after the context is destroyed. */
typedef struct qx_snap qx_snap;
enum { QX_SNAP_PSI = 1, QX_SNAP_V = 2 };
qx_status qx_snap_create(qx_ctx *ctx, int what, qx_snap **out);
qx_status qx_snap_take(qx_snap *snap);
qx_status qx_snap_wait(qx_snap *snap, double *psi_c128, double *V);
qx_status qx_snap_destroy(qx_snap *snap);

/* ---- split-step operators ------------------------------------------------ */
/* Field::kick (src/field.cxx:93-116): psi <- psi exp(-i dt V) */
qx_status qx_kick(qx_ctx *ctx, double dt);
/* Field::drift (src/field.cxx:119-198): psi <- IFFT(e^{-i dt k^2/2}/S FFT psi) */
qx_status qx_drift(qx_ctx *ctx, double dt);
/* Field::updatePotential (src/field.cxx:294-346); vmax nullable */
qx_status qx_update_potential(qx_ctx *ctx, double *vmax);
/* Field::Poisson (src/field.cxx:230-291) applied to a caller half spectrum
 * (reference layout, complex128, modified in place) */
qx_status qx_poisson_half(qx_ctx *ctx, double *vhat_c128);
/* Field::computeEnergy (src/field.cxx:200-227): V <- |psi|^2, rho_max/mean */
qx_status qx_compute_energy(qx_ctx *ctx, double *rhomax_over_mean);
/* density statistics that leave V untouched; outputs nullable.  mass =
 * (L/N)^dim sum |psi|^2. */
qx_status qx_density_stats(qx_ctx *ctx, double *rhomax_over_mean, double *mass);
/* qx_density_stats without allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  psi: a device array of S = N^dim complex
 * values of the given precision (row-major, as a context's psi); the
 * statistics go to out[0] = rho_max/mean and out[1] = mass = (L/N)^dim
 * sum |psi|^2, a caller-owned device array of two doubles, equal bit for
 * bit to qx_density_stats on the same psi.  Stream-ordered on `stream` (a
 * hipStream_t; null = the default stream); returns without waiting, and
 * allocates nothing: the reduction's partials go to a static 16 KiB device
 * array (1024 pairs of doubles), one per GPU, held until the process ends.
 * Lattices as qx_create accepts them; other sizes are rejected
 * (QX_EINVAL).  Every call on one GPU shares that GPU's array, so calls
 * that overlap on one GPU (different streams, or threads) race. */
qx_status qx_density_stats_noalloc(const void *psi_dev, int dim, int N, double L, qx_precision prec,
                                   void *stream, double *out_dev);
/* mass, kinetic 1/2 dV/S sum k^2|psihat|^2 and potential 1/2 dV sum V|psi|^2
 * (no reference counterpart; outputs nullable) */
qx_status qx_energy(qx_ctx *ctx, double *mass, double *kinetic, double *potential);
/* Vmax of the last potential update (Field::Vmax, src/field.h:56) */
double qx_vmax(const qx_ctx *ctx);

/* Spectrum::compute (src/spectrum/spectrum.cxx:32-164).  source 0: the real
 * field V as it stands (the reference's choice); 1: |psi|^2 computed on the
 * fly, leaving V untouched.  nbins/spb: bin j gathers the integer shells
 * [j*spb, (j+1)*spb); the reference is nbins = N/2+1, spb = 1.  Outputs
 * (arrays of nbins): shell-averaged P, mode counts, k of the bin. */
qx_status qx_spectrum(qx_ctx *ctx, int source, int nbins, int spb, double *Pk, int64_t *count,
                      double *kmodes);

/* ---- step compositions ---------------------------------------------------
 * One merged-kick leapfrog segment of the reference time loop
 * (src/main.cxx:65-85): requires V current (after qx_update_potential).
 * Opens with K(ds/2) if first_half else K(ds) (continuing a segment left
 * open), then nsteps x [D(ds), P] with K(ds) between steps, and closes with
 * K(ds/2) if last_half.  Leaves V current.  Runs the fused pass schedule. */
qx_status qx_leapfrog(qx_ctx *ctx, double ds, int nsteps, int first_half, int last_half);
/* The same loop with constants that change from step to step: the MRE run
 * of the reference, whose Field::updateTime (src/field.cxx:349-361) sets a
 * new ds and a after every drift, so that pref = norm_ a (:298) changes
 * too.  Requires V current.  K(kick_dt[0]); then for step j < nsteps:
 * D(drift_dt[j]), P with pref[j] (the context's rho_bar and k = 0 mode),
 * K(kick_dt[j+1]); kick_dt[nsteps] = 0 leaves the segment open.  Arrays:
 * drift_dt[nsteps], kick_dt[nsteps+1], pref[nsteps].  The context's
 * physics are not changed.  Fused pass schedule, one step's tables per
 * launch set (not graph-captured). */
qx_status qx_leapfrog_schedule(qx_ctx *ctx, int nsteps, const double *drift_dt, const double *kick_dt,
                               const double *pref);
/* Field::propagate (src/field.cxx:388-402): K(ds/2) D(ds) P K(ds/2) */
qx_status qx_propagate(qx_ctx *ctx, double ds);

/* ---- FFTBackend (src/fft/fft.h:9-26), host arrays ---------------------- */
/* forward_c2c / backward_c2c (src/fft/fft.h:122-130): in place, S complex,
 * sign -1 forward, +1 backward, unnormalised */
qx_status qx_fft_c2c(qx_ctx *ctx, double *data_c128, int sign);
/* forward_r2c (src/fft/fft.h:132-135): S real -> H complex half spectrum */
qx_status qx_fft_r2c(qx_ctx *ctx, const double *in, double *out_c128);
/* forward_r2c of the context's own V, as Field::updatePotential
 * (src/field.cxx:312) and Spectrum::compute (src/spectrum/spectrum.cxx:48)
 * call it on the Field's V: H complex half spectrum (reference layout),
 * times `scale` (1/S gives the Poisson-multiplied spectrum that the last
 * potential update transformed back into V) */
qx_status qx_fft_r2c_V(qx_ctx *ctx, double scale, double *out_c128);
/* backward_c2r (src/fft/fft.h:137-140): H complex -> S real; `in` is not
 * modified (FFTW may destroy it, so callers cannot rely on either) */
qx_status qx_fft_c2r(qx_ctx *ctx, const double *in_c128, double *out);
/* sites() / ksites() / global_N() (src/fft/fft.h:19-25) */
int64_t qx_sites(const qx_ctx *ctx);
int64_t qx_ksites(const qx_ctx *ctx);
int qx_global_N(const qx_ctx *ctx);
/* lattice dimension and state precision of a context (no reference counterpart) */
int qx_dim(const qx_ctx *ctx);
qx_precision qx_precision_of(const qx_ctx *ctx);

/* ---- initial conditions on the device (no reference counterpart except
 * qx_ic_solitons, SolitonsIC::apply src/ic/solitons.cxx:20-110) ----------- */
/* solitons: rows (amp, sigma, x, y, z) in box fractions, background bkg */
qx_status qx_ic_solitons(qx_ctx *ctx, const double *sol, int nsol, double bkg);
/* Gaussian random field: delta with P(k) ~ k^-2 for 0 < |n| <= N/4,
 * psi = sqrt(max(0, 1 + amp delta/sigma)) e^{i theta}, theta = U[0,2pi)
 * noise low-passed at the same cutoff; counter-based RNG keyed by seed */
qx_status qx_ic_grf(qx_ctx *ctx, uint64_t seed, double amp);
/* blobs: rows (rho_c, r_c, x, y, z, phi), background density bkg */
qx_status qx_ic_blobs(qx_ctx *ctx, const double *table, int nblobs, double bkg);

/* ---- measurement (no reference counterpart; replaces the reference's
 * PROFILE/FLOP_COUNT section timers, src/profiler.h:158-185) -------------- */
/* passes of one fused steady-state KDK step (7 in 3D, 4 in 2D) */
int qx_step_pass_count(const qx_ctx *ctx);
const char *qx_step_pass_name(const qx_ctx *ctx, int pass);
/* Times each pass inside `reps` consecutive fused steps (after one warm-up
 * step) on scratch copies of the state (the state itself, drift counter
 * included, is left untouched), with HIP events between the passes on the
 * context's stream.
 * ms[pass] = average ms per launch; bytes[pass] (nullable) = minimum HBM
 * bytes the pass moves.  cap >= pass count. */
qx_status qx_time_step_passes(qx_ctx *ctx, double ds, int reps, double *ms, double *bytes, int cap);

/* ---- ensembles of independent 2D lattices (no reference counterpart: the
 * reference runs one lattice per process) ---------------------------------
 * Members may have different N (power of two in [16, 4096]) and L; members
 * of one N form a size class whose state is stored member after member, so
 * each pass of the step covers the whole class in one launch.  Steady-state
 * steps are captured in a hipGraph (graph_steps steps per replay, size
 * classes as parallel branches).  Physics constants are shared; ds is per
 * member.  Each member evolves exactly as it would alone in a qx_ctx. */
typedef struct qx_ens qx_ens;
qx_status qx_ens_create(int nmembers, const int *N, const double *L, qx_precision prec, int device,
                        qx_ens **out);
qx_status qx_ens_destroy(qx_ens *ens);
qx_status qx_ens_set_physics(qx_ens *ens, const qx_physics *ph);
/* steps per captured graph; 0 launches every pass directly */
qx_status qx_ens_set_graph_steps(qx_ens *ens, int steps);
qx_status qx_ens_upload_psi(qx_ens *ens, int member, const double *psi_c128);
qx_status qx_ens_download_psi(qx_ens *ens, int member, double *psi_c128);
/* per-member GRF, identical to qx_ic_grf on a lone lattice */
qx_status qx_ens_ic_grf(qx_ens *ens, const uint64_t *seeds, double amp);
qx_status qx_ens_update_potential(qx_ens *ens);
/* qx_leapfrog for every member at once; ds[member]; nsteps >= 1 */
qx_status qx_ens_leapfrog(qx_ens *ens, const double *ds, int nsteps, int first_half, int last_half);
/* Autotuner for ensembles: per size class, the faster of the two tile
 * variants of its strided drift pass, timed over `reps` steady steps of the
 * class on saved copies of its state (the state is restored).  variants
 * (nullable): one int per class, classes in increasing N */
qx_status qx_ens_tune(qx_ens *ens, const double *ds, int reps, int *variants);
/* Reinstall a tuning result (one variant per class, as qx_ens_tune returns
 * them: 0 full tile, 1 half tile), e.g. one kept from an earlier run.  The
 * choice changes speed, never results. */
qx_status qx_ens_set_tile_variants(qx_ens *ens, const int *variants);
int qx_ens_classes(const qx_ens *ens);
int qx_ens_class_size(const qx_ens *ens, int cls);
/* mass[member] = (L/N)^2 sum |psi|^2 */
qx_status qx_ens_mass(qx_ens *ens, double *mass);
int qx_ens_size(const qx_ens *ens);
int64_t qx_ens_sites(const qx_ens *ens);
void *qx_ens_stream(qx_ens *ens);

/* ---- learned coarse-step corrector (no reference counterpart) -------------
 * Periodic 3x3 convolutions 3 -> 32 -> 32 -> 32 -> 2 with GELU (erf form)
 * on (Re psi, Im psi, a V) of an H x W lattice: psi <- psi + s net(.).  BF16
 * weights and hidden activations, FP32 accumulation; the 32 -> 32 forward
 * layers run on MFMA.  Parameters: 19970 floats, per layer W[co][ci][dy][dx]
 * then b[co], tap (dy, dx) reading (h + dy - 1, w + dx - 1) mod (H, W).
 * Input device/host arrays x are [H*W][3] float (Re psi, Im psi, a V);
 * outputs and targets are [H*W][2] float (a complex64 psi array). */
typedef struct qx_net qx_net;
int qx_net_param_count(void);
qx_status qx_net_create(int H, int W, int device, qx_net **out);
qx_status qx_net_destroy(qx_net *net);
qx_status qx_net_set_params(qx_net *net, const float *theta);  /* resets Adam state */
/* fixed output scale s: y = x + s net(x) (default 1); set it to the typical
 * size of the correction so the network works with O(1) outputs */
qx_status qx_net_set_output_scale(qx_net *net, float s);
float qx_net_output_scale(const qx_net *net);
/* a: the V channel's scale where the input is built from a lattice
 * (correct, coarse_steps, train_step_lattices); default 1.  a = 0 feeds the
 * channel zeros: the network is then a function of psi alone, the 2-channel
 * network of the 2 -> 32 -> 32 -> 32 -> 2 layout (the V channel's weights
 * see no input and get no gradient) */
qx_status qx_net_set_input_scale(qx_net *net, float a);
float qx_net_input_scale(const qx_net *net);
qx_status qx_net_get_params(qx_net *net, float *theta);
qx_status qx_net_forward(qx_net *net, const void *x_dev, void *y_dev);     /* y = x[:, 0:2] + s net(x) */
qx_status qx_net_forward_host(qx_net *net, const float *x, float *y);
/* loss = mean((x[:, 0:2] + s net(x) - target)^2) and its gradient (grad_host nullable) */
qx_status qx_net_loss_grad(qx_net *net, const void *x_dev, const void *t_dev, double *loss, float *grad_host);
qx_status qx_net_loss_grad_host(qx_net *net, const float *x, const float *t, double *loss, float *grad);
/* Adam on the last gradient (bias-corrected) */
qx_status qx_net_adam_step(qx_net *net, float lr, float beta1, float beta2, float eps);
qx_status qx_net_train_step(qx_net *net, const void *x_dev, const void *t_dev, float lr, double *loss);
/* one Adam step on the pair (input from coarse's psi and V, target fine's psi) */
qx_status qx_net_train_step_lattices(qx_net *net, qx_ctx *coarse, qx_ctx *fine, float lr, double *loss);
/* mean((x - target)^2) over two psi arrays [H*W][2]: the loss an identity
 * corrector would have */
qx_status qx_net_pair_mse(qx_net *net, const void *x_dev, const void *t_dev, double *mse);
/* psi <- psi + s net(psi, a V) on a 2D FP32 context of size H x W, then
 * scaled back to the mass it had before (V untouched) */
qx_status qx_net_correct(qx_net *net, qx_ctx *ctx);
/* energy gate of qx_net_coarse_steps: a chunk whose relative change of the
 * energy (kinetic + potential, qx_energy) exceeds tol is redone with fine
 * steps, as one beyond mass_tol is; 0 (the default) turns the gate off */
qx_status qx_net_set_energy_tol(qx_net *net, double tol);
/* nsteps corrected coarse steps of k*dt each (leapfrog(k dt, 1), correct,
 * scale psi back to its mass at the start of the call, update the
 * potential), stream-ordered with no host waits inside a chunk of
 * check_every steps; after each chunk the relative mass change (and the
 * energy's, with an energy tolerance set) is read back and, beyond its
 * tolerance, the chunk is redone as k*chunk fine steps from the saved
 * state (*fallbacks = number of redone chunks, nullable) */
qx_status qx_net_coarse_steps(qx_net *net, qx_ctx *ctx, double dt, int k, int nsteps, int check_every,
                              double mass_tol, int *fallbacks);
void *qx_net_stream(qx_net *net);
const char *qx_net_last_error(void);

/* Autotuner: times both tile variants (full / half column tile) of every
 * strided pass of the fused step on scratch copies, twice in alternation,
 * and keeps the faster per pass kind.  variants (nullable, 8 ints): the choice per [op kind][stride
 * class], op kinds forward, inverse, drift, Poisson. */
qx_status qx_tune(qx_ctx *ctx, double ds, int reps, int *variants);
/* installs variants recorded by an earlier qx_tune (e.g. from a cache) */
qx_status qx_set_tile_variants(qx_ctx *ctx, const int *variants);
/* Streaming (non-temporal) loads and stores in the step's passes (no
 * reference counterpart: a cache policy, results are bitwise the same).
 * qx_create turns them on for lattices whose psi is at least 512 MB; on = 1
 * or 0 overrides that.  Lattices with N < 512 have no streaming kernels and
 * report 0.  *on_out (nullable) receives the setting in force. */
qx_status qx_set_stream_io(qx_ctx *ctx, int on, int *on_out);

/* ---- plumbing (no reference counterpart) -------------------------------- */
void *qx_stream(qx_ctx *ctx);          /* the context's hipStream_t */
qx_status qx_synchronize(qx_ctx *ctx);
const char *qx_last_error(void);
const char *qx_version(void);

#ifdef __cplusplus
}
#endif
#endif
