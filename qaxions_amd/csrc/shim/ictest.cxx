// ictest.cxx -- the reference's pybind11 IC test module
// (src/examples/ictest.cxx:8-50): a 2D or 3D Field from defaults() (norm 1)
// with the soliton IC from ./solitons.txt; Re psi at idx.
#include <pybind11/pybind11.h>
#include "field.h"
#include "ic.h"

namespace {
double soliton_psi(int dim, int N, int idx)
{
    Params pars;
    defaults(pars);
    pars.N = N;
    pars.dim = dim;
    pars.norm = 1;
    pars.verb = false;
    Field field(pars);
    auto ic = createIC(pars.ictype, pars);
    ic->apply(field);
    return field.getPsiIdx(idx);
}
}  // namespace

double test_solitons(int N, int idx) { return soliton_psi(2, N, idx); }     // src/examples/ictest.cxx:8-24
double test_solitons_3D(int N, int idx) { return soliton_psi(3, N, idx); }  // src/examples/ictest.cxx:26-42

PYBIND11_MODULE(ictest, m)
{
    m.def("test_solitons", &test_solitons, "");
    m.def("test_solitons_3D", &test_solitons_3D, "");
}
