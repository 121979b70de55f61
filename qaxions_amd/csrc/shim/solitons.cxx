// solitons.cxx -- SolitonsIC (src/ic/solitons.cxx:20-110) and createIC
// (src/ic/ic.cxx:4-13).  x is the fastest lattice index; centres are box
// fractions; there are no periodic images; rho is normalised to mean 1 and
// psi = sqrt(rho) is real.
#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <vector>
#include "ic.h"

namespace {
struct Sol {
    double amp, sigma, norm, x, y, z;
};
}  // namespace

void SolitonsIC::apply(Field &field) const
{
    const int N = field.size(), dim = field.dim();
    const double L = field.Lbox();
    const size_t sites = field.sites();
    const double dx = L / N;
    std::vector<Sol> sols;
    std::ifstream file("solitons.txt");
    if (!file.is_open()) throw std::runtime_error("Error opening file solitons.txt !");
    std::string line;
    while (std::getline(file, line)) {
        std::istringstream ss(line);
        Sol s;
        if (ss >> s.amp >> s.sigma >> s.x >> s.y >> s.z) {
            s.norm = 1 / (s.sigma * s.sigma * s.sigma * 2 * M_PI);
            sols.push_back(s);
        } else if (!line.empty()) {
            std::cerr << "Error parsing line: " << line << std::endl;
        }
    }
    std::vector<double> rho(sites, p_.sol_bkg);
    double sum = 0.0;
    for (size_t idx = 0; idx < sites; ++idx) {
        double x, y, z = 0.0;
        if (dim == 3) {
            x = (double)(idx % N) * dx;
            y = (double)((idx / N) % N) * dx;
            z = (double)(idx / ((size_t)N * N)) * dx;
        } else {
            x = (double)(idx % N) * dx;
            y = (double)(idx / N) * dx;
        }
        for (const Sol &s : sols) {
            const double ddx = x - s.x * L, ddy = y - s.y * L;
            const double ddz = (dim == 3) ? (z - s.z * L) : 0.0;
            const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            rho[idx] += s.amp * s.norm * std::exp(-r2 / (2 * s.sigma * s.sigma));
        }
        sum += rho[idx];
    }
    const double mean = sum / static_cast<double>(sites);
    qx_complex *psi = field.psi();
    for (size_t idx = 0; idx < sites; ++idx) {
        psi[idx][0] = std::sqrt(rho[idx] / mean);
        psi[idx][1] = 0.0;
    }
}

std::unique_ptr<InitialCondition> createIC(ICType type, const Params &p)
{
    switch (type) {
    case ICType::SOLITONS: return std::make_unique<SolitonsIC>(p);
    default: throw std::runtime_error("Unknown ICType");
    }
}
