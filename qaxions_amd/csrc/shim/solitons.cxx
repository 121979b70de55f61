// solitons.cxx -- the soliton initial condition over the device Field
// (SolitonsIC, src/ic/solitons.cxx:20-110) and the factory createIC
// (src/ic/ic.cxx:4-13).  The host reads the table; the field is formed on
// the GPU by qx_ic_solitons: Gaussians amp/(2 pi sigma^3) exp(-r^2/(2 sigma^2))
// on the background, the mean density normalised to 1, psi = sqrt(rho) real.
#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>
#include "ic.h"

namespace {

constexpr int COLUMNS = 5;  // amplitude, width, centre x, y, z (box fractions)

// One soliton per line: the first five numbers of a line.  Blank lines are
// skipped; any other line without five numbers is reported and skipped,
// which is how the reference treats it.
std::vector<double> parse_soliton_table(const std::string &text, const std::string &source)
{
    std::vector<double> table;
    size_t begin = 0;
    for (int lineno = 1; begin < text.size(); ++lineno) {
        size_t end = text.find('\n', begin);
        if (end == std::string::npos) end = text.size();
        const std::string line = text.substr(begin, end - begin);
        begin = end + 1;
        double row[COLUMNS];
        int have = 0;
        const char *cur = line.c_str();
        while (have < COLUMNS) {
            char *next = nullptr;
            errno = 0;
            const double x = std::strtod(cur, &next);
            if (next == cur || errno != 0) break;
            row[have++] = x;
            cur = next;
        }
        if (have == COLUMNS)
            table.insert(table.end(), row, row + COLUMNS);
        else if (line.find_first_not_of(" \t\r") != std::string::npos)
            std::cerr << source << ":" << lineno << ": not a soliton (" << COLUMNS
                      << " numbers expected), skipped: " << line << std::endl;
    }
    return table;
}

}  // namespace

void SolitonsIC::apply(Field &field) const
{
    // the table is read from the working directory (src/ic/solitons.cxx:32)
    const std::string source = "solitons.txt";
    std::ifstream in(source, std::ios::binary);
    if (!in) throw std::runtime_error("SolitonsIC: cannot open " + source);
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    const std::vector<double> table = parse_soliton_table(text, source);
    field.beginDeviceWrite();
    qx_throw(qx_ic_solitons(field.context(), table.empty() ? nullptr : table.data(),
                            (int)(table.size() / COLUMNS), p_.sol_bkg));
    field.endDeviceWrite();
}

std::unique_ptr<InitialCondition> createIC(ICType type, const Params &p)
{
    if (type == ICType::SOLITONS) return std::make_unique<SolitonsIC>(p);
    throw std::runtime_error("createIC: no initial condition of this type");
}
