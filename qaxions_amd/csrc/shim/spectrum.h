// spectrum.h -- the reference's power spectrum (src/spectrum/spectrum.h:17-56).
// compute() bins the spectrum of V as it stands on the device, and leaves
// the Field's Vhat holding r2c(V) (src/spectrum/spectrum.cxx:48).
#ifndef QX_SHIM_SPECTRUM_H
#define QX_SHIM_SPECTRUM_H

#include <string>
#include <vector>

class Field;

class Spectrum {
  public:
    explicit Spectrum(Field &field);
    // reference binning: N/2+1 integer shells
    void compute();
    // bins of `spb` consecutive shells (the first nbins of them)
    void compute(int nbins, int spb);
    // the same binning of |psi|^2's spectrum, formed on the device without
    // touching V or Vhat (no reference counterpart: measure() without the
    // reference's clobbering of V)
    void computeDensity();
    void computeDensity(int nbins, int spb);
    void write(const std::string &filename) const;
    const std::vector<double> &Pk() const { return Pk_; }
    const std::vector<int> &counts() const { return count_; }
    const std::vector<double> &kmodes() const { return k_; }

  private:
    void run(int source, int nbins, int spb);
    Field &f_;
    int N_;
    double a_;
    std::vector<double> Pk_;
    std::vector<int> count_;
    std::vector<double> k_;
};

#endif
