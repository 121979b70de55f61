// field.cxx -- Field over the qx C ABI.  Each method cites the reference
// method whose effect it reproduces.
#include "field.h"

#include <cmath>
#include <iostream>
#include <vector>

void defaults(Params &p)  // src/parse.cxx:7-29
{
    p.N = 64;
    p.dim = 3;
    p.nthr = 1;
    p.ai = 0.1;
    p.Lbox = 1.0;
    p.dtr = 4.0;
    p.dt = 0.0001;
    p.nsteps = 10;
    p.nmeas = 20;
    p.norm = 4000;
    p.sol_bkg = 1.0;
    p.verb = false;
    p.readj = false;
    p.dir = "output";
    p.ictype = ICType::SOLITONS;
    p.cosmotype = CosmoType::STATIC;
    p.plan = FFTPlanType::ESTIMATE;
    p.measinfo = MeasureType::NONE;
    p.precision = Precision::FP64;
    p.device = 0;
    p.k0_mode = 1;
    p.rho_bar_mode = 0;
}

Field::Field() {}

Field::Field(Params &p) : Field() { init(p); }

Field::~Field() { releaseContext(); }

void Field::releaseContext()
{
    fft_backend_.reset();
    if (ctx_) qx_destroy(ctx_);
    ctx_ = nullptr;
}

// src/field.cxx:13-46
void Field::init(Params &p)
{
    releaseContext();
    N_ = p.N;
    Lbox_ = p.Lbox;
    dim_ = p.dim;
    nthr_ = p.nthr;
    nsteps_ = p.nsteps;
    verb_ = p.verb;
    k0_mode_ = p.k0_mode;
    rho_bar_mode_ = p.rho_bar_mode;
    setCosmo(p);
    if (dim_ != 2 && dim_ != 3) throw std::invalid_argument("Dimension must be 2 or 3");
    qx_throw(qx_create(dim_, N_, Lbox_, p.precision == Precision::FP64 ? QX_FP64 : QX_FP32,
                       p.device, &ctx_));
    fft_backend_ = std::make_unique<QxBackend>(ctx_);
    sites_ = fft_backend_->sites();
    ksites_ = fft_backend_->ksites();
    psi_h_.assign(2 * sites_, 0.0);
    V_h_.assign(sites_, 0.0);
    Vhat_.assign(2 * ksites_, 0.0);
    vhat_src_ = VhatSrc::Host;
    psi_host_valid_ = V_host_valid_ = true;
    psi_host_dirty_ = V_host_dirty_ = false;
    setPhysics();
}

// src/field.cxx:68-91
void Field::setCosmo(Params &p)
{
    switch (p.cosmotype) {
    case CosmoType::STATIC:
        cosmo_ = 0;
        a_ = 1.0;
        s_ = 0;
        ds_ = p.dt;
        rho_mean_ = 1.0;
        norm_ = 4 * M_PI * p.norm;
        break;
    case CosmoType::MRE:
        cosmo_ = 1;
        a_ = p.ai;
        s_ = 0;
        ds_ = p.dt / a_;
        rho_mean_ = 1.5 * (2.0 * M_PI) * (2.0 * M_PI) * 1.27 * 1.27;
        norm_ = 1.0;
        break;
    }
}

void Field::setPhysics()
{
    if (!ctx_) return;
    qx_physics ph{norm_ * a_, rho_mean_, k0_mode_, rho_bar_mode_};  // pref: src/field.cxx:298
    qx_throw(qx_set_physics(ctx_, &ph));
}

void Field::syncToDevice() const
{
    if (psi_host_dirty_) {
        qx_throw(qx_upload_psi(ctx_, psi_h_.data()));
        psi_host_dirty_ = false;
        psi_host_valid_ = true;
    }
    if (V_host_dirty_) {
        qx_throw(qx_upload_V(ctx_, V_h_.data()));
        V_host_dirty_ = false;
        V_host_valid_ = true;
    }
}

const qx_complex *Field::psi() const
{
    if (!psi_host_valid_) {
        qx_throw(qx_download_psi(ctx_, psi_h_.data()));
        psi_host_valid_ = true;
    }
    return (const qx_complex *)psi_h_.data();
}
qx_complex *Field::psi()
{
    const_cast<const Field *>(this)->psi();
    psi_host_dirty_ = true;
    return (qx_complex *)psi_h_.data();
}
const double *Field::V() const
{
    if (!V_host_valid_) {
        qx_throw(qx_download_V(ctx_, V_h_.data()));
        V_host_valid_ = true;
    }
    return V_h_.data();
}
double *Field::V()
{
    const_cast<const Field *>(this)->V();
    beforeVChange();
    V_host_dirty_ = true;
    return V_h_.data();
}

qx_complex *Field::Vhat()
{
    if (vhat_src_ != VhatSrc::Host) {
        syncToDevice();
        const double scale = vhat_src_ == VhatSrc::PoissonOfV ? 1.0 / (double)sites_ : 1.0;
        qx_throw(qx_fft_r2c_V(ctx_, scale, Vhat_.data()));
        vhat_src_ = VhatSrc::Host;
    }
    return (qx_complex *)Vhat_.data();
}

void Field::beforeVChange() const
{
    if (vhat_src_ == VhatSrc::R2cOfV) const_cast<Field *>(this)->Vhat();
    vhat_src_ = VhatSrc::Host;
}

void Field::beginDeviceWrite()
{
    syncToDevice();
    beforeVChange();
}

void Field::endDeviceWrite()
{
    psi_host_valid_ = false;
    V_host_valid_ = false;
}

// src/field.cxx:93-116
void Field::kick(double dt)
{
    if (verb_) std::cout << "[kick] Starting loop ..." << std::endl;
    syncToDevice();
    qx_throw(qx_kick(ctx_, dt));
    psi_host_valid_ = false;
    if (verb_) std::cout << "[kick] done!" << std::endl;
}

// src/field.cxx:119-198
void Field::drift(double dt)
{
    syncToDevice();
    qx_throw(qx_drift(ctx_, dt));
    psi_host_valid_ = false;
}

// src/field.cxx:230-291: Poisson factor applied to Vhat() in place
void Field::Poisson()
{
    double *vh = (double *)Vhat();
    setPhysics();
    qx_throw(qx_poisson_half(ctx_, vh));
}

// src/field.cxx:294-346
void Field::updatePotential()
{
    syncToDevice();
    setPhysics();
    double vmax = 0.0;
    qx_throw(qx_update_potential(ctx_, &vmax));
    Vmax_ = vmax;
    V_host_valid_ = false;
    vhat_src_ = VhatSrc::PoissonOfV;
}

// src/field.cxx:200-227: overwrites V with |psi|^2 and sets rhomax
void Field::computeEnergy()
{
    syncToDevice();
    beforeVChange();
    double r = 0.0;
    qx_throw(qx_compute_energy(ctx_, &r));
    rhomax_ = r;
    V_host_valid_ = false;
}

// src/field.cxx:349-361
void Field::updateTime()
{
    curr_ += 1;
    if (cosmo_ == 1) {
        ds_ = 0.00025 / a_;
        a_ = a_ * (1 + 1.27 * 1.27 * std::sqrt(1 + a_) * ds_);
    }
    s_ += ds_;
}

void Field::half_kick() { kick(0.5 * ds_); }  // src/field.cxx:363-368
void Field::full_kick() { kick(ds_); }        // src/field.cxx:370-376

void Field::drift_update()  // src/field.cxx:378-386
{
    drift(ds_);
    updateTime();
    updatePotential();
}

void Field::propagate()  // src/field.cxx:388-402
{
    kick(0.5 * ds_);
    drift(ds_);
    updateTime();
    updatePotential();
    kick(0.5 * ds_);
}

void Field::leapfrog(int nsteps, bool first_half, bool last_half)
{
    if (cosmo_ != 0) {
        if (nsteps == 0) {
            if (first_half) half_kick(); else full_kick();
            if (last_half) half_kick();
            return;
        }
        // ds and a change after every drift (updateTime): unroll the
        // schedule on the host and run it as one fused loop
        std::vector<double> drift_dt(nsteps), kick_dt(nsteps + 1), pref(nsteps);
        kick_dt[0] = first_half ? 0.5 * ds_ : ds_;
        for (int i = 0; i < nsteps; ++i) {
            drift_dt[i] = ds_;
            updateTime();
            pref[i] = norm_ * a_;
            kick_dt[i + 1] = i != nsteps - 1 ? ds_ : (last_half ? 0.5 * ds_ : 0.0);
        }
        syncToDevice();
        setPhysics();
        qx_throw(qx_leapfrog_schedule(ctx_, nsteps, drift_dt.data(), kick_dt.data(), pref.data()));
        Vmax_ = qx_vmax(ctx_);
        psi_host_valid_ = false;
        V_host_valid_ = false;
        vhat_src_ = VhatSrc::PoissonOfV;
        return;
    }
    syncToDevice();
    setPhysics();
    qx_throw(qx_leapfrog(ctx_, ds_, nsteps, first_half ? 1 : 0, last_half ? 1 : 0));
    for (int i = 0; i < nsteps; ++i) updateTime();
    Vmax_ = qx_vmax(ctx_);
    psi_host_valid_ = false;
    if (nsteps > 0) {
        V_host_valid_ = false;
        vhat_src_ = VhatSrc::PoissonOfV;
    }
}
