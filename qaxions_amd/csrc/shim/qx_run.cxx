// qx_run.cxx -- the reference's program (src/main.cxx:21-90) over the
// device Field: soliton IC from ./solitons.txt, measure() at the steps of
// generateMeasList, the time loop of half_kick / drift_update / full_kick.
// Between two measurements the loop is one Field::leapfrog segment (the
// merged kicks fused into the step passes); a segment ends after a step's
// potential update with the merged kick pending, where the reference
// measures.  Grid files are written in the background by a SnapshotWriter.
// The final state is written with psi to <dir>/final.h5.
//
// The reference parses a command line (src/parse.cxx); this program takes
// positional arguments instead:  qx_run [N [dim [nsteps [nmeas [dir [fp32]]]]]]
#include <cstdlib>
#include <iostream>
#include <string>
#include "field.h"
#include "ic.h"
#include "io.h"
#include "meas.h"

int main(int argc, char **argv)
{
    try {
        Params pars;
        defaults(pars);
        pars.verb = false;
        pars.norm = 1;
        if (argc > 1) pars.N = std::atoi(argv[1]);
        if (argc > 2) pars.dim = std::atoi(argv[2]);
        if (argc > 3) pars.nsteps = std::atoi(argv[3]);
        if (argc > 4) pars.nmeas = std::atoi(argv[4]);
        if (argc > 5) pars.dir = argv[5];
        if (argc > 6 && std::string(argv[6]) == "fp32") pars.precision = Precision::FP32;
        pars.measinfo = MeasureType::SPECTRUM | MeasureType::RHO_GRID;
        if (pars.nsteps < 1) throw std::invalid_argument("nsteps must be >= 1");

        const Clock::time_point start = Clock::now();
        Field field(pars);
        auto ic = createIC(pars.ictype, pars);
        ic->apply(field);

        const size_t nsteps = (size_t)pars.nsteps;
        const std::vector<size_t> mlist = generateMeasList(nsteps, (size_t)pars.nmeas);
        SnapshotWriter writer(field, false);  // rho/data = V after computeEnergy, as the reference
        size_t measn = 0, next = 0;
        measure(field, pars, 0, start, true, &writer);
        field.updatePotential();

        size_t done = 0;  // steps taken
        bool first = true;
        while (done < nsteps) {
            // a repeated entry of the list (more measurements than steps) is
            // never matched again by the reference's loop, which then stops
            // measuring (src/main.cxx:75-79)
            if (next < mlist.size() && mlist[next] < done) next = mlist.size();
            // run up to and including the next measured step (or to the end)
            const bool measured = next < mlist.size();
            const size_t upto = measured ? mlist[next] + 1 : nsteps;
            const bool end = upto == nsteps;
            field.leapfrog((int)(upto - done), first, end && !measured);
            first = false;
            done = upto;
            if (measured) {
                ++measn;
                ++next;
                measure(field, pars, measn, start, true, &writer);
                if (end) field.half_kick();  // the final half kick follows the last measurement
            }
        }
        writer.wait();
        IO io(pars.dir + "/final.h5", FileMode::Create);
        io.writeConf(field, true);
        io.close();
        std::cout << "qx_run: " << nsteps << " steps, " << measn + 1 << " measurements, done in "
                  << formatDuration(Clock::now() - start) << std::endl;
        return 0;
    } catch (const std::exception &e) {
        std::cerr << "qx_run: " << e.what() << std::endl;
        return 1;
    }
}
