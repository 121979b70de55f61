// meas.cxx -- measure() (src/meas.cxx:13-63), including the HDF5 grid
// output (RHO_GRID, src/meas.cxx:48-59).
#include "meas.h"

#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include "io.h"
#include "spectrum.h"

void measure(Field &f, const Params &p, size_t measNumber, Clock::time_point st, bool legacy_clobber,
             SnapshotWriter *writer)
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
        if (legacy_clobber)
            spec.compute();          // V holds |psi|^2 after computeEnergy
        else
            spec.computeDensity();   // |psi|^2 formed on the device, V untouched
        std::ostringstream fname;
        fname << p.dir << "/spec_" << std::setw(4) << std::setfill('0') << measNumber << ".txt";
        spec.write(fname.str());
    }
    if (meas & MeasureType::RHO_GRID) {
        if (f.verb()) std::cout << "[measure] Saving energy grid" << std::endl;
        std::ostringstream fname;
        fname << p.dir << "/field_" << std::setw(4) << std::setfill('0') << measNumber << ".h5";
        if (writer) {
            if (writer->rhoFromPsi() == legacy_clobber)
                throw std::invalid_argument("measure: the writer's rho_from_psi must be !legacy_clobber");
            writer->write(f, fname.str());
        } else if (legacy_clobber) {
            IO io(fname.str());  // V holds |psi|^2 after computeEnergy
            io.writeConf(f, false);
            io.close();
        } else {
            const qx_complex *psi = f.psi();
            std::vector<double> rho(f.sites());
            for (size_t i = 0; i < rho.size(); ++i) rho[i] = psi[i][0] * psi[i][0] + psi[i][1] * psi[i][1];
            IO io(fname.str());
            io.writeConfArrays(ConfHeader{f.size(), f.dim(), f.curr(), f.time(), f.a(), f.Lbox()}, rho.data(),
                               nullptr);
            io.close();
        }
    }
    if (f.verb()) std::cout << "\n\n";
}
