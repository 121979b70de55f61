// fft.h -- the reference's FFT plugin interface (src/fft/fft.h:9-26) and its
// GPU implementation.  Arrays are host arrays, as with FFTWOpenMPBackend;
// each call moves them through the device and back.
#ifndef QX_SHIM_FFT_H
#define QX_SHIM_FFT_H

#include <cstddef>
#include <stdexcept>
#include <string>
#include "qx.h"

typedef double qx_complex[2];  // layout-compatible with fftw_complex

class FFTBackend {
  public:
    virtual ~FFTBackend() = default;
    virtual void forward_c2c(qx_complex *data) = 0;
    virtual void backward_c2c(qx_complex *data) = 0;
    virtual void forward_r2c(double *in_real, qx_complex *out_complex) = 0;
    virtual void backward_c2r(qx_complex *in_complex, double *out_real) = 0;
    virtual size_t sites() const = 0;
    virtual size_t ksites() const = 0;
    virtual bool owns_zero_mode() const { return true; }
    virtual bool owns_zero_mode_k() const { return true; }
    virtual ptrdiff_t global_N() const = 0;
};

// Turns a qx status into the exception the reference raises for that case.
inline void qx_throw(qx_status s)
{
    if (s == QX_OK) return;
    std::string msg = qx_last_error();
    if (s == QX_EINVAL) throw std::invalid_argument(msg);
    if (s == QX_ENOMEM) throw std::bad_alloc();
    throw std::runtime_error(msg);
}

// FFTBackend over a qx context (replaces FFTWOpenMPBackend, src/fft/fft.h:28-149)
class QxBackend : public FFTBackend {
  public:
    explicit QxBackend(qx_ctx *ctx) : ctx_(ctx) {}
    void forward_c2c(qx_complex *data) override { qx_throw(qx_fft_c2c(ctx_, &data[0][0], -1)); }
    void backward_c2c(qx_complex *data) override { qx_throw(qx_fft_c2c(ctx_, &data[0][0], +1)); }
    void forward_r2c(double *in, qx_complex *out) override { qx_throw(qx_fft_r2c(ctx_, in, &out[0][0])); }
    void backward_c2r(qx_complex *in, double *out) override { qx_throw(qx_fft_c2r(ctx_, &in[0][0], out)); }
    size_t sites() const override { return (size_t)qx_sites(ctx_); }
    size_t ksites() const override { return (size_t)qx_ksites(ctx_); }
    ptrdiff_t global_N() const override { return (ptrdiff_t)qx_global_N(ctx_); }

  private:
    qx_ctx *ctx_;
};

#endif
