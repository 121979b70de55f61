// meas.h -- measurement dispatcher and schedule (src/meas.h:28-29,
// src/utils.h), restated over the device Field.
#ifndef QX_SHIM_MEAS_H
#define QX_SHIM_MEAS_H

#include <chrono>
#include <string>
#include <vector>
#include "field.h"

using Clock = std::chrono::steady_clock;

std::vector<size_t> generateMeasList(size_t num_steps, size_t n_bins);
std::string formatDuration(std::chrono::duration<double> dur);
void printStatus(const Field &f, int step, size_t num_steps, int meas, std::chrono::duration<double> dur);

class SnapshotWriter;

// src/meas.cxx:13-63.  legacy_clobber = true keeps the reference's
// side effect of computeEnergy (V <- |psi|^2, which the next kick then
// uses); false computes the statistics without touching V, and RHO_GRID
// files then hold |psi|^2 formed from psi.  With a writer (made with
// rho_from_psi = !legacy_clobber) the RHO_GRID file <dir>/field_NNNN.h5 is
// written in the background; otherwise before measure returns.  Built
// into libqxio.so with the HDF5 IO.
void measure(Field &f, const Params &p, size_t measNumber, Clock::time_point st,
             bool legacy_clobber = true, SnapshotWriter *writer = nullptr);

#endif
