// field.h -- the reference's Field (src/field.h:13-97) with its state on the
// GPU.  Public methods keep the reference's names, arguments and effects;
// psi()/V() hand out host mirrors that are synchronised lazily: a mutable
// accessor marks the mirror dirty and the next operator uploads it, and a
// device-side update is downloaded on the next access.
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

    qx_complex *Vhat() { return Vhat_.empty() ? nullptr : (qx_complex *)Vhat_.data(); }

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

  private:
    void setPhysics();
    void releaseContext();

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
    std::vector<double> Vhat_;
    mutable bool psi_host_valid_ = true, psi_host_dirty_ = false;
    mutable bool V_host_valid_ = true, V_host_dirty_ = false;
};

#endif
