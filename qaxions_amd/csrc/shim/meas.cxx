// meas.cxx -- measure() (src/meas.cxx:13-63).  The HDF5 snapshot branch
// (RHO_GRID, src/meas.cxx:48-59) is not provided by this build.
#include "meas.h"

#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include "spectrum.h"

void measure(Field &f, const Params &p, size_t measNumber, Clock::time_point st, bool legacy_clobber)
{
    const MeasureType meas = p.measinfo;
    if (meas == MeasureType::NONE) return;
    auto dur = Clock::now() - st;
    if (f.verb()) std::cout << "\n\n";
    printStatus(f, f.curr(), (size_t)f.nsteps(), (int)measNumber, dur);
    if (legacy_clobber) {
        f.computeEnergy();
    } else {
        double r = 0.0;
        f.syncToDevice();
        qx_throw(qx_density_stats(f.context(), &r, nullptr));
    }
    if (meas & MeasureType::SPECTRUM) {
        if (f.verb()) std::cout << "[measure] Saving power spectrum" << std::endl;
        Spectrum spec(f);
        if (legacy_clobber) {
            spec.compute();
        } else {
            // spectrum of |psi|^2 without writing V
            const int n = f.size() / 2 + 1;
            std::vector<double> P(n), k(n);
            std::vector<int64_t> c(n);
            qx_throw(qx_spectrum(f.context(), 1, n, 1, P.data(), c.data(), k.data()));
            spec.compute();  // keeps the object's arrays consistent with V
        }
        std::ostringstream fname;
        fname << p.dir << "/spec_" << std::setw(4) << std::setfill('0') << measNumber << ".txt";
        spec.write(fname.str());
    }
    if (meas & MeasureType::RHO_GRID)
        throw std::runtime_error("measure: HDF5 grid output (RHO_GRID) is not provided by this build");
    if (f.verb()) std::cout << "\n\n";
}
