// fieldtest.cxx -- the reference's pybind11 test module
// (src/examples/fieldtest.cxx:11-105): each entry point builds a Field from
// defaults() with the soliton IC, runs one operator sequence and returns one
// real scalar.  `fp32` (keyword, default False) selects complex64 device state.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include "field.h"
#include "ic.h"
#include "spectrum.h"

#include <complex>

namespace py = pybind11;

namespace {
std::unique_ptr<Field> make_field(int n, int N, bool fp32, bool mre = false)
{
    Params pars;
    defaults(pars);
    pars.N = N;
    pars.dim = n;
    pars.norm = 1;
    pars.verb = false;
    pars.precision = fp32 ? Precision::FP32 : Precision::FP64;
    if (mre) pars.cosmotype = CosmoType::MRE;
    auto field = std::make_unique<Field>(pars);
    auto ic = createIC(pars.ictype, pars);
    ic->apply(*field);
    return field;
}
}  // namespace

double test_kick(int n, int N, double dt, int idx, bool fp32)
{
    auto field = make_field(n, N, fp32);
    field->updatePotential();
    field->kick(dt);
    return field->getPsiIdx(idx);
}

double test_drift(int n, int N, double dt, int idx, bool fp32)
{
    auto field = make_field(n, N, fp32);
    field->drift(dt);
    return field->getPsiIdx(idx);
}

double test_potential(int n, int N, int idx, bool fp32)
{
    auto field = make_field(n, N, fp32);
    field->updatePotential();
    return field->getVIdx(idx);
}

double test_kick_drift_step(int n, int N, double dt, int idx, bool fp32)
{
    auto field = make_field(n, N, fp32);
    field->updatePotential();
    field->kick(0.5 * dt);
    field->drift(dt);
    field->updatePotential();
    field->kick(0.5 * dt);
    return field->getPsiIdx(idx);
}

// Not in the reference module: an MRE run of the time loop
// (src/main.cxx:65-85) from the soliton IC, either as Field::leapfrog's one
// fused call (fused) or operator by operator as the reference's loop calls
// them.  Returns (psi, a, ds, s) after nsteps.
py::tuple mre_leapfrog(int n, int N, int nsteps, bool fused, bool fp32)
{
    auto field = make_field(n, N, fp32, true);
    field->updatePotential();
    if (fused) {
        field->leapfrog(nsteps, true, true);
    } else {
        field->half_kick();
        for (int i = 0; i < nsteps; ++i) {
            field->drift_update();
            if (i != nsteps - 1) field->full_kick();
        }
        field->half_kick();
    }
    const size_t S = field->sites();
    py::array_t<std::complex<double>> out(S);
    const qx_complex *p = field->psi();
    auto *o = out.mutable_data();
    for (size_t i = 0; i < S; ++i) o[i] = std::complex<double>(p[i][0], p[i][1]);
    return py::make_tuple(out, field->a(), field->ds(), field->time());
}

namespace {
py::array_t<std::complex<double>> complex_array(const qx_complex *p, size_t n)
{
    py::array_t<std::complex<double>> out(n);
    auto *o = out.mutable_data();
    for (size_t i = 0; i < n; ++i) o[i] = std::complex<double>(p[i][0], p[i][1]);
    return out;
}
}  // namespace

// Not in the reference module: the Field's Vhat() through the reference's
// sequences (soliton IC): after updatePotential (the Poisson spectrum the
// potential was formed from); after computeEnergy and Spectrum::compute
// (r2c of V = |psi|^2, src/spectrum/spectrum.cxx:48); then after
// Field::Poisson() applied to that spectrum in place.  Returns (psi, the
// three half spectra, V after updatePotential).
py::tuple vhat_sequence(int n, int N, bool fp32)
{
    auto field = make_field(n, N, fp32);
    const size_t S = field->sites(), H = field->ksites();
    auto psi = complex_array(static_cast<const Field &>(*field).psi(), S);
    field->updatePotential();
    py::array_t<double> V(S);
    const double *v = static_cast<const Field &>(*field).V();
    std::copy(v, v + S, V.mutable_data());
    auto after_potential = complex_array(field->Vhat(), H);
    field->computeEnergy();
    Spectrum spec(*field);
    spec.compute();
    auto after_spectrum = complex_array(field->Vhat(), H);
    field->Poisson();
    auto after_poisson = complex_array(field->Vhat(), H);
    return py::make_tuple(psi, after_potential, after_spectrum, after_poisson, V);
}

PYBIND11_MODULE(fieldtest, m)
{
    m.def("vhat_sequence", &vhat_sequence, "", py::arg("n"), py::arg("N"), py::arg("fp32") = false);
    m.def("test_kick", &test_kick, "", py::arg("n"), py::arg("N"), py::arg("dt"), py::arg("idx"),
          py::arg("fp32") = false);
    m.def("test_drift", &test_drift, "", py::arg("n"), py::arg("N"), py::arg("dt"), py::arg("idx"),
          py::arg("fp32") = false);
    m.def("test_potential", &test_potential, "", py::arg("n"), py::arg("N"), py::arg("idx"),
          py::arg("fp32") = false);
    m.def("test_kick_drift_step", &test_kick_drift_step, "", py::arg("n"), py::arg("N"),
          py::arg("dt"), py::arg("idx"), py::arg("fp32") = false);
    m.def("mre_leapfrog", &mre_leapfrog, "", py::arg("n"), py::arg("N"), py::arg("nsteps"),
          py::arg("fused") = true, py::arg("fp32") = false);
}
