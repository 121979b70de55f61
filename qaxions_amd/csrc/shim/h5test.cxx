// h5test.cxx -- the reference's pybind11 HDF5 test module
// (src/examples/h5test.cxx:11-72): test_read_write and test_read_jaxions
// keep their names, arguments and results.  The other entry points are
// additions: host-array forms of the file formats (no device needed) and
// device runs of the background writer, restart and measure().
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include "field.h"
#include "ic.h"
#include "io.h"
#include "meas.h"

#include <algorithm>
#include <cmath>
#include <complex>

namespace py = pybind11;
using cplx = std::complex<double>;

namespace {

Params soliton_params(int n, int N, bool fp32)
{
    Params pars;
    defaults(pars);
    pars.verb = false;
    if (n > 0) {
        pars.dim = n;
        pars.N = N;
        pars.norm = 1;
    }
    pars.precision = fp32 ? Precision::FP32 : Precision::FP64;
    return pars;
}

py::array_t<cplx> psi_array(const Field &f)
{
    py::array_t<cplx> out(f.sites());
    const qx_complex *p = f.psi();
    std::copy((const cplx *)&p[0][0], (const cplx *)&p[0][0] + f.sites(), out.mutable_data());
    return out;
}

py::array_t<double> real_array(const double *v, size_t n)
{
    py::array_t<double> out(n);
    std::copy(v, v + n, out.mutable_data());
    return out;
}

py::dict header_dict(const ConfHeader &h)
{
    py::dict d;
    d["N"] = h.N;
    d["dim"] = h.dim;
    d["step"] = h.step;
    d["time"] = h.time;
    d["a"] = h.a;
    d["Lbox"] = h.Lbox;
    return d;
}

}  // namespace

// src/examples/h5test.cxx:11-50: write a soliton Field with psi, read it
// into a new Field, return the larger difference of Re/Im psi at idx
double test_read_write(int idx, const std::string &fname)
{
    Params pars = soliton_params(0, 0, false);
    auto field1 = std::make_unique<Field>(pars);
    auto ic = createIC(pars.ictype, pars);
    ic->apply(*field1);
    const double re1 = field1->getPsiIdx(idx), im1 = field1->getPsiImagIdx(idx);
    {
        IO writer(fname, FileMode::Create);
        writer.writeConf(*field1, true);
        writer.flush();
    }
    Field field2;
    Params pars2 = pars;
    {
        IO reader(fname, FileMode::ReadOnly);
        reader.readConf(pars2, field2);
    }
    return std::max(std::abs(re1 - field2.getPsiIdx(idx)), std::abs(im1 - field2.getPsiImagIdx(idx)));
}

// src/examples/h5test.cxx:52-65: Re psi[idx] of ./jaxions_2D.hdf5
double test_read_jaxions(int idx)
{
    Params pars = soliton_params(0, 0, false);
    auto field = std::make_unique<Field>();
    IO reader("jaxions_2D.hdf5", FileMode::ReadOnly);
    reader.readJaxions(pars, *field);
    return field->getPsiIdx(idx);
}

// ---- host-array forms -----------------------------------------------------------------
void write_conf_arrays(const std::string &fname, int N, int dim, int step, double time, double a, double Lbox,
                       py::array_t<double, py::array::c_style | py::array::forcecast> rho, py::object psi)
{
    const ConfHeader h{N, dim, step, time, a, Lbox};
    size_t S = 1;
    for (int d = 0; d < dim; ++d) S *= (size_t)N;
    if ((size_t)rho.size() != S) throw std::invalid_argument("rho must hold N^dim values");
    py::array_t<cplx, py::array::c_style | py::array::forcecast> p;
    if (!psi.is_none()) {
        p = psi.cast<py::array_t<cplx, py::array::c_style | py::array::forcecast>>();
        if ((size_t)p.size() != S) throw std::invalid_argument("psi must hold N^dim values");
    }
    IO io(fname, FileMode::Create);
    io.writeConfArrays(h, rho.data(), psi.is_none() ? nullptr : (const double *)p.data());
}

py::tuple read_conf_arrays(const std::string &fname, bool with_psi)
{
    IO io(fname, FileMode::ReadOnly);
    std::vector<double> psi;
    const ConfHeader h = io.readConfArrays(with_psi ? &psi : nullptr);
    if (!with_psi) return py::make_tuple(header_dict(h), py::none());
    py::array_t<cplx> out(psi.size() / 2);
    std::copy((const cplx *)psi.data(), (const cplx *)psi.data() + psi.size() / 2, out.mutable_data());
    return py::make_tuple(header_dict(h), out);
}

py::array_t<double> read_rho(const std::string &fname)
{
    IO io(fname, FileMode::ReadOnly);
    std::vector<double> rho;
    io.readConfArrays(nullptr, &rho);
    return real_array(rho.data(), rho.size());
}

void write_jaxions_arrays(const std::string &fname, int N, int dim, double L,
                          py::array_t<double, py::array::c_style | py::array::forcecast> m,
                          py::array_t<double, py::array::c_style | py::array::forcecast> v)
{
    size_t S = 1;
    for (int d = 0; d < dim; ++d) S *= (size_t)N;
    if ((size_t)m.size() != S || (size_t)v.size() != S) throw std::invalid_argument("m and v must hold N^dim values");
    IO io(fname, FileMode::Create);
    io.writeJaxionsArrays(N, dim, L, m.data(), v.data());
}

py::tuple read_jaxions_arrays(const std::string &fname)
{
    IO io(fname, FileMode::ReadOnly);
    std::vector<double> psi;
    const ConfHeader h = io.readJaxionsArrays(&psi);
    py::array_t<cplx> out(psi.size() / 2);
    std::copy((const cplx *)psi.data(), (const cplx *)psi.data() + psi.size() / 2, out.mutable_data());
    return py::make_tuple(h.N, h.dim, h.Lbox, out);
}

// ---- device runs -------------------------------------------------------------------------

// A STATIC soliton run of `nsteps` fused steps with a background snapshot
// after `snap_at` of them (positions at integer time, the merged kick still
// pending, as where the reference's loop measures), then a restart from
// the file: readConf, updatePotential and the remaining steps with the
// kick opening the segment.  Returns the state downloaded just before the
// snapshot, the file's contents, and both runs' final psi.
py::dict snapshot_restart(int n, int N, int nsteps, int snap_at, const std::string &fname, bool fp32)
{
    Params pars = soliton_params(n, N, fp32);
    Field a(pars);
    createIC(pars.ictype, pars)->apply(a);
    a.updatePotential();
    a.leapfrog(snap_at, true, false);
    py::dict out;
    out["psi_at_snapshot"] = psi_array(a);
    out["V_at_snapshot"] = real_array(static_cast<const Field &>(a).V(), a.sites());
    {
        SnapshotWriter w(a, true);
        w.write(a, fname);                       // returns at once
        a.leapfrog(nsteps - snap_at, false, true);  // the loop runs on
        w.wait();
    }
    out["psi_final"] = psi_array(a);
    Field b;
    Params pb = pars;
    {
        IO io(fname, FileMode::ReadOnly);
        io.readConf(pb, b);
    }
    b.updatePotential();
    b.leapfrog(nsteps - snap_at, false, true);
    out["psi_restart_final"] = psi_array(b);
    return out;
}

// measure() with SPECTRUM | RHO_GRID into `dir` after one fused step;
// returns the psi it measured
py::array_t<cplx> measure_run(int n, int N, const std::string &dir, int meas_number, bool legacy_clobber,
                              bool background, bool fp32)
{
    Params pars = soliton_params(n, N, fp32);
    pars.dir = dir;
    pars.measinfo = MeasureType::SPECTRUM | MeasureType::RHO_GRID;
    Field f(pars);
    createIC(pars.ictype, pars)->apply(f);
    f.updatePotential();
    f.leapfrog(1, true, false);
    auto psi = psi_array(f);
    if (background) {
        SnapshotWriter w(f, false, !legacy_clobber);
        measure(f, pars, (size_t)meas_number, Clock::now(), legacy_clobber, &w);
        w.wait();
    } else {
        measure(f, pars, (size_t)meas_number, Clock::now(), legacy_clobber);
    }
    return psi;
}

PYBIND11_MODULE(h5test, m)
{
    m.def("test_read_write", &test_read_write, "");
    m.def("test_read_jaxions", &test_read_jaxions, "");
    m.def("write_conf_arrays", &write_conf_arrays, py::arg("fname"), py::arg("N"), py::arg("dim"),
          py::arg("step"), py::arg("time"), py::arg("a"), py::arg("Lbox"), py::arg("rho"),
          py::arg("psi") = py::none());
    m.def("read_conf_arrays", &read_conf_arrays, py::arg("fname"), py::arg("with_psi") = true);
    m.def("read_rho", &read_rho);
    m.def("write_jaxions_arrays", &write_jaxions_arrays);
    m.def("read_jaxions_arrays", &read_jaxions_arrays);
    m.def("snapshot_restart", &snapshot_restart, py::arg("n"), py::arg("N"), py::arg("nsteps"),
          py::arg("snap_at"), py::arg("fname"), py::arg("fp32") = false);
    m.def("measure_run", &measure_run, py::arg("n"), py::arg("N"), py::arg("dir"), py::arg("meas_number"),
          py::arg("legacy_clobber") = true, py::arg("background") = false, py::arg("fp32") = false);
}
