// field.h -- the reference's Field (src/field.h:13-97) with its state on the
// GPU.  Public methods keep the reference's names, arguments and effects;
// psi()/V()/Vhat() hand out host mirrors that are synchronised lazily: a
// mutable accessor of psi or V marks the mirror dirty and the next operator
// uploads it, and a device-side update is downloaded on the next access.
//
// Vhat holds what the reference's Vhat_ holds: r2c(V) after
// Spectrum::compute (src/spectrum/spectrum.cxx:48), and after
// updatePotential the Poisson-multiplied spectrum that was transformed back
// into V (src/field.cxx:312-332; the reference's buffer is then whatever
// FFTW's c2r left in its input, which is not specified, so here it holds the
// spectrum itself).  Both are formed from the device's V when Vhat() is
// read: a spectrum of V is downloaded before anything rewrites V, the
// Poisson spectrum is dropped then (the reference has none to keep).
#ifndef QX_SHIM_FIELD_H
#define QX_SHIM_FIELD_H

#include <memory>
#include <vector>
#include "fft.h"
#include "parse.h"

class Field {
  public:
    void init(Params &pars);
    void setCosmo(Params &pars);

    Field();
    explicit Field(Params &pars);
    ~Field();
    Field(const Field &) = delete;
    Field &operator=(const Field &) = delete;

    void kick(double dt);
    void drift(double dt);
    void Poisson();
    void updatePotential();
    void computeEnergy();
    void updateTime();
    void half_kick();
    void full_kick();
    void drift_update();
    void propagate();
    // nsteps of the merged-kick loop (src/main.cxx:69-85) in one fused call
    void leapfrog(int nsteps, bool first_half, bool last_half);

    const qx_complex *psi() const;
    qx_complex *psi();
    double getPsiIdx(int idx) const { return psi()[idx][0]; }
    double getPsiImagIdx(int idx) const { return psi()[idx][1]; }

    const double *V() const;
    double *V();
    double getVIdx(int idx) const { return V()[idx]; }

    qx_complex *Vhat();

    FFTBackend *fftBackend() const { return fft_backend_.get(); }

    int size() const { return N_; }
    int dim() const { return dim_; }
    double Lbox() const { return Lbox_; }
    size_t sites() const { return sites_; }
    size_t ksites() const { return ksites_; }
    double Vmax() const { return Vmax_; }
    double rhomax() const { return rhomax_; }
    double rho_mean() const { return rho_mean_; }
    bool verb() const { return verb_; }
    int curr() const { return curr_; }
    double time() const { return s_; }
    double a() const { return a_; }
    double ds() const { return ds_; }
    int nsteps() const { return nsteps_; }
    int cosmo() const { return cosmo_; }

    qx_ctx *context() const { return ctx_; }
    // push host edits to the device (called by every device operator)
    void syncToDevice() const;
    // plumbing for the shim's own callers (no reference counterpart):
    // Spectrum::compute has transformed V, so Vhat is r2c(V)
    void vhatIsSpectrumOfV() { vhat_src_ = VhatSrc::R2cOfV; }
    // a device-side writer (an initial condition) is about to rewrite psi
    // and V: pending host edits go first, the host mirrors are refreshed on
    // the next access
    void beginDeviceWrite();
    void endDeviceWrite();

  private:
    void setPhysics();
    void releaseContext();
    // called before the device's V is rewritten (see Vhat above)
    void beforeVChange() const;

    int N_ = 0;
    int dim_ = 0;
    int nthr_ = 0;
    double Lbox_ = 0.0;
    double norm_ = 0.0;
    double rho_mean_ = 0.0;
    bool verb_ = false;
    int curr_ = 0;
    int nsteps_ = 0;
    int cosmo_ = 0;
    double ds_ = 0.0;
    double s_ = 0.0;
    double a_ = 1.0;
    size_t sites_ = 0;
    size_t ksites_ = 0;
    double Vmax_ = 0.0;
    double rhomax_ = 0.0;
    int k0_mode_ = 1;
    int rho_bar_mode_ = 0;

    qx_ctx *ctx_ = nullptr;
    std::unique_ptr<FFTBackend> fft_backend_;
    // host mirrors
    mutable std::vector<double> psi_h_, V_h_;
    mutable std::vector<double> Vhat_;
    // where the current Vhat lives: the host array, or to be formed from the
    // device's V as r2c(V) or as r2c(V)/S (the Poisson spectrum)
    enum class VhatSrc { Host, R2cOfV, PoissonOfV };
    mutable VhatSrc vhat_src_ = VhatSrc::Host;
    mutable bool psi_host_valid_ = true, psi_host_dirty_ = false;
    mutable bool V_host_valid_ = true, V_host_dirty_ = false;
};

#endif
