// utils.cxx -- measurement schedule and status line (src/utils.cxx:8-85)
#include <iomanip>
#include <iostream>
#include <sstream>
#include "meas.h"

std::vector<size_t> generateMeasList(size_t num_steps, size_t n_bins)  // src/utils.cxx:8-28
{
    std::vector<size_t> bins;
    bins.reserve(n_bins);
    if (n_bins == 0 || num_steps == 0) return bins;
    if (n_bins == 1) {
        bins.push_back(num_steps - 1);
        return bins;
    }
    for (size_t k = 0; k < n_bins; ++k) bins.push_back(k * (num_steps - 1) / (n_bins - 1));
    return bins;
}

std::string formatDuration(std::chrono::duration<double> dur)  // src/utils.cxx:30-62
{
    const double seconds = dur.count();
    std::ostringstream out;
    out << std::fixed << std::setprecision(1);
    if (seconds < 1.0) {
        out << static_cast<int>(seconds * 1000) << "ms";
    } else if (seconds < 60.0) {
        out << seconds << "s";
    } else if (seconds < 3600.0) {
        const int mins = static_cast<int>(seconds) / 60;
        out << mins << "m " << seconds - mins * 60 << "s";
    } else {
        const int hrs = static_cast<int>(seconds) / 3600;
        const int mins = (static_cast<int>(seconds) % 3600) / 60;
        out << hrs << "h " << mins << "m " << seconds - hrs * 3600 - mins * 60 << "s";
    }
    return out.str();
}

void printStatus(const Field &f, int step, size_t num_steps, int meas,
                 std::chrono::duration<double> dur)  // src/utils.cxx:64-85
{
    const double percent = 100.0 * step / (double)num_steps;
    const std::string bar(98, '=');
    std::cout << bar << "\n";
    std::cout << std::fixed << std::setprecision(1) << std::setw(6) << "Meas #" << std::setw(4) << meas
              << std::setw(2) << " |   a: " << std::setw(8) << std::setprecision(4) << f.a()
              << std::setw(2) << " |   max(δ): " << std::setw(12) << std::setprecision(5) << f.rhomax()
              << std::setw(4) << "    |  " << std::setw(6) << std::setprecision(1) << percent << "%"
              << std::setw(12) << "Walltime:" << " " << std::setw(12) << formatDuration(dur) << std::endl;
    std::cout << bar << "\n";
}
