// spectrum.cxx -- Spectrum over qx_spectrum (src/spectrum/spectrum.cxx:11-183)
#include "spectrum.h"

#include <fstream>
#include <stdexcept>
#include "field.h"

Spectrum::Spectrum(Field &f) : f_(f), N_(f.size()), a_(f.a())
{
    if (!f.context()) throw std::runtime_error("PowerSpectrum: Field buffers are null.");
    const int n = N_ / 2 + 1;
    Pk_.assign(n, 0.0);
    count_.assign(n, 0);
    k_.assign(n, 0.0);
}

void Spectrum::compute() { compute(N_ / 2 + 1, 1); }

void Spectrum::compute(int nbins, int spb)
{
    run(0, nbins, spb);
    f_.vhatIsSpectrumOfV();
}

void Spectrum::computeDensity() { computeDensity(N_ / 2 + 1, 1); }

void Spectrum::computeDensity(int nbins, int spb) { run(1, nbins, spb); }

void Spectrum::run(int source, int nbins, int spb)
{
    f_.syncToDevice();
    std::vector<int64_t> cnt(nbins);
    Pk_.assign(nbins, 0.0);
    k_.assign(nbins, 0.0);
    qx_throw(qx_spectrum(f_.context(), source, nbins, spb, Pk_.data(), cnt.data(), k_.data()));
    count_.assign(cnt.begin(), cnt.end());
}

// ASCII "k P count" for non-empty shells (src/spectrum/spectrum.cxx:167-183)
void Spectrum::write(const std::string &filename) const
{
    std::ofstream out(filename);
    if (!out) throw std::runtime_error("Spectrum::write: cannot open " + filename);
    out << "# a: " << a_ << std::endl;
    out << "# k_phys      P(k)          count\n";
    for (int n = 0; n < static_cast<int>(Pk_.size()); ++n) {
        if (count_[n] == 0) continue;
        out << std::scientific << k_[n] << "  " << Pk_[n] << "  " << count_[n] << "\n";
    }
}
