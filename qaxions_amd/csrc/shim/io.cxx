// io.cxx -- HDF5 configuration files over the HDF5 C library.  Each IO
// method cites the reference method whose file format and effect it
// reproduces (src/io/io.cxx).  HDF5 errors become std::runtime_error, as
// in the reference.
#include "io.h"

#include <cmath>
#include <complex>
#include <stdexcept>
#include <type_traits>

namespace {

// an hid_t closed on scope exit
struct Hid {
    hid_t id;
    herr_t (*closer)(hid_t);
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), closer(c) {}
    ~Hid()
    {
        if (id >= 0) closer(id);
    }
    Hid(const Hid &) = delete;
    Hid &operator=(const Hid &) = delete;
    bool ok() const { return id >= 0; }
};

template <typename T>
hid_t native_type()
{
    if constexpr (std::is_same_v<T, int>) return H5T_NATIVE_INT;
    else if constexpr (std::is_same_v<T, double>) return H5T_NATIVE_DOUBLE;
    else if constexpr (std::is_same_v<T, float>) return H5T_NATIVE_FLOAT;
    else if constexpr (std::is_same_v<T, unsigned int>) return H5T_NATIVE_UINT;
    else static_assert(!sizeof(T), "unsupported attribute type");
}

size_t sites_of(int N, int dim)
{
    size_t S = 1;
    for (int d = 0; d < dim; ++d) S *= (size_t)N;
    return S;
}

// the Field's complex128 psi as interleaved doubles
std::vector<double> psi_copy(const Field &f)
{
    const qx_complex *p = f.psi();
    return std::vector<double>(&p[0][0], &p[0][0] + 2 * f.sites());
}

}  // namespace

// src/io/io.cxx:14-43
IO::IO(const std::string &filename, FileMode mode)
{
    switch (mode) {
    case FileMode::Create: file_ = H5Fcreate(filename.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT); break;
    case FileMode::ReadOnly: file_ = H5Fopen(filename.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT); break;
    case FileMode::ReadWrite: file_ = H5Fopen(filename.c_str(), H5F_ACC_RDWR, H5P_DEFAULT); break;
    }
    if (file_ < 0) {
        const char *what = mode == FileMode::Create     ? "create/truncate"
                           : mode == FileMode::ReadOnly ? "open (read-only)"
                                                        : "open (read-write)";
        throw std::runtime_error(std::string("Failed to ") + what + " HDF5 file: " + filename);
    }
}

IO::~IO() { close(); }

void IO::flush()  // src/io/io.cxx:46-51
{
    if (file_ >= 0) H5Fflush(file_, H5F_SCOPE_GLOBAL);
}

void IO::close()  // src/io/io.cxx:54-61
{
    if (file_ >= 0) {
        H5Fflush(file_, H5F_SCOPE_GLOBAL);
        H5Fclose(file_);
        file_ = -1;
    }
}

void IO::requireOpen() const
{
    if (file_ < 0) throw std::runtime_error("File not open");
}

void IO::ensureGroup(const std::string &name)  // src/io/io.cxx:169-186
{
    requireOpen();
    const htri_t e = H5Lexists(file_, name.c_str(), H5P_DEFAULT);
    if (e > 0) return;
    if (e < 0) throw std::runtime_error("H5Lexists failed for group: " + name);
    Hid g(H5Gcreate2(file_, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) throw std::runtime_error("Failed to create group: " + name);
}

bool IO::hasAttribute(const std::string &group, const std::string &name) const
{
    Hid g(H5Gopen2(file_, group.c_str(), H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return false;
    return H5Aexists(g.id, name.c_str()) > 0;
}

// src/io/io.cxx:70-104
template <typename T>
T IO::readAttribute(const std::string &group, const std::string &name) const
{
    requireOpen();
    Hid g(H5Gopen2(file_, group.c_str(), H5P_DEFAULT), H5Gclose);
    if (!g.ok()) throw std::runtime_error("Cannot open group: " + group);
    Hid a(H5Aopen(g.id, name.c_str(), H5P_DEFAULT), H5Aclose);
    if (!a.ok()) throw std::runtime_error("Attribute '" + name + "' not found in group: " + group);
    T v{};
    if (H5Aread(a.id, native_type<T>(), &v) < 0)
        throw std::runtime_error("Failed to read attribute '" + name + "' in group: " + group);
    return v;
}

// src/io/io.cxx:106-166: a 1-element rank-1 attribute, the group created if absent
template <typename T>
void IO::writeAttribute(const std::string &group, const std::string &name, const T &value)
{
    requireOpen();
    ensureGroup(group);
    Hid g(H5Gopen2(file_, group.c_str(), H5P_DEFAULT), H5Gclose);
    if (!g.ok()) throw std::runtime_error("Cannot open group: " + group);
    const hsize_t one = 1;
    Hid sp(H5Screate_simple(1, &one, nullptr), H5Sclose);
    if (!sp.ok()) throw std::runtime_error("Failed to create attribute dataspace");
    Hid a(H5Acreate2(g.id, name.c_str(), native_type<T>(), sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (!a.ok()) throw std::runtime_error("Failed to create attribute: " + name);
    if (H5Awrite(a.id, native_type<T>(), &value) < 0) throw std::runtime_error("Failed to write attribute: " + name);
}

// src/io/io.cxx:188-207, with the element count checked against the Field
void IO::readDataset(const std::string &path, double *out, size_t count)
{
    requireOpen();
    Hid d(H5Dopen2(file_, path.c_str(), H5P_DEFAULT), H5Dclose);
    if (!d.ok()) throw std::runtime_error("Cannot open dataset: " + path);
    Hid sp(H5Dget_space(d.id), H5Sclose);
    const hssize_t n = H5Sget_simple_extent_npoints(sp.id);
    if (n < 0 || (size_t)n != count)
        throw std::runtime_error("Dataset " + path + " holds " + std::to_string((long long)n) +
                                 " elements, expected " + std::to_string(count));
    if (H5Dread(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
        throw std::runtime_error("Failed to read dataset: " + path);
}

// src/io/io.cxx:209-235: dims {N,N,N} in 3D, {N,N,1} in 2D (src/io/io.cxx:325-328)
void IO::writeDataset(const std::string &group, const std::string &name, int N, int dim, const double *data)
{
    requireOpen();
    ensureGroup(group);
    const std::string path = group + "/" + name;
    const hsize_t dims[3] = {(hsize_t)N, (hsize_t)N, dim == 3 ? (hsize_t)N : (hsize_t)1};
    Hid sp(H5Screate_simple(3, dims, nullptr), H5Sclose);
    if (!sp.ok()) throw std::runtime_error("Failed to create dataspace for: " + path);
    Hid d(H5Dcreate2(file_, path.c_str(), H5T_NATIVE_DOUBLE, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
          H5Dclose);
    if (!d.ok()) throw std::runtime_error("Failed to create dataset: " + path);
    if (H5Dwrite(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
        throw std::runtime_error("Failed to write dataset: " + path);
}

// ---- configuration files ----------------------------------------------------------

void IO::writeConfArrays(const ConfHeader &h, const double *rho, const double *psi_c128)
{
    if (h.dim != 2 && h.dim != 3) throw std::invalid_argument("dim must be 2 or 3");
    const size_t S = sites_of(h.N, h.dim);
    writeDataset("rho", "data", h.N, h.dim, rho);
    if (psi_c128) {
        std::vector<double> part(S);
        for (int c = 0; c < 2; ++c) {
            for (size_t i = 0; i < S; ++i) part[i] = psi_c128[2 * i + c];
            writeDataset("psi", c == 0 ? "real" : "imag", h.N, h.dim, part.data());
        }
    }
    writeAttribute<int>("Header", "N", h.N);
    writeAttribute<int>("Header", "dim", h.dim);
    writeAttribute<int>("Header", "step", h.step);
    writeAttribute<double>("Header", "time", h.time);
    writeAttribute<double>("Header", "a", h.a);
    writeAttribute<double>("Header", "Lbox", h.Lbox);
}

ConfHeader IO::readConfArrays(std::vector<double> *psi_c128, std::vector<double> *rho)
{
    ConfHeader h;
    h.N = readAttribute<int>("Header", "N");
    h.dim = readAttribute<int>("Header", "dim");
    h.Lbox = readAttribute<double>("Header", "Lbox");
    if (hasAttribute("Header", "step")) h.step = readAttribute<int>("Header", "step");
    if (hasAttribute("Header", "time")) h.time = readAttribute<double>("Header", "time");
    if (hasAttribute("Header", "a")) h.a = readAttribute<double>("Header", "a");
    if (h.dim != 2 && h.dim != 3) throw std::runtime_error("Header dim must be 2 or 3");
    if (psi_c128) {
        const size_t S = sites_of(h.N, h.dim);
        std::vector<double> part(S);
        psi_c128->assign(2 * S, 0.0);
        for (int c = 0; c < 2; ++c) {
            readDataset(c == 0 ? "/psi/real" : "/psi/imag", part.data(), S);
            for (size_t i = 0; i < S; ++i) (*psi_c128)[2 * i + c] = part[i];
        }
    }
    if (rho) {
        rho->resize(sites_of(h.N, h.dim));
        readDataset("/rho/data", rho->data(), rho->size());
    }
    return h;
}

ConfHeader IO::readJaxionsArrays(std::vector<double> *psi_c128)
{
    ConfHeader h;
    h.N = readAttribute<int>("Header", "N");
    h.dim = readAttribute<int>("Header", "dim");
    h.Lbox = readAttribute<double>("Header", "L");
    if (h.dim != 2 && h.dim != 3) throw std::runtime_error("Header dim must be 2 or 3");
    if (psi_c128) {
        const size_t S = sites_of(h.N, h.dim);
        std::vector<double> m(S), v(S);
        readDataset("/m", m.data(), S);
        readDataset("/v", v.data(), S);
        // the reference's fixed initial scale factor and its 12-digit sqrt(2)
        // (src/io/io.cxx:12,264-265)
        const double ai = 0.01, scale = std::pow(ai, 1.5) / 1.41421356237;
        psi_c128->resize(2 * S);
        for (size_t i = 0; i < S; ++i) {
            (*psi_c128)[2 * i] = scale * m[i];
            (*psi_c128)[2 * i + 1] = -scale * v[i];
        }
    }
    return h;
}

void IO::writeJaxionsArrays(int N, int dim, double L, const double *m, const double *v)
{
    if (dim != 2 && dim != 3) throw std::invalid_argument("dim must be 2 or 3");
    const size_t S = sites_of(N, dim);
    const hsize_t dims[3] = {(hsize_t)N, (hsize_t)N, dim == 3 ? (hsize_t)N : (hsize_t)1};
    for (int k = 0; k < 2; ++k) {
        Hid sp(H5Screate_simple(3, dims, nullptr), H5Sclose);
        Hid d(H5Dcreate2(file_, k == 0 ? "/m" : "/v", H5T_NATIVE_DOUBLE, sp.id, H5P_DEFAULT, H5P_DEFAULT,
                         H5P_DEFAULT),
              H5Dclose);
        if (!sp.ok() || !d.ok() || H5Dwrite(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, k == 0 ? m : v) < 0)
            throw std::runtime_error("Failed to write jaxions dataset");
    }
    (void)S;
    writeAttribute<int>("Header", "N", N);
    writeAttribute<int>("Header", "dim", dim);
    writeAttribute<double>("Header", "L", L);
}

// src/io/io.cxx:276-311
void IO::readConf(Params &pars, Field &field)
{
    requireOpen();
    std::vector<double> psi;
    const ConfHeader h = readConfArrays(&psi);
    pars.N = h.N;
    pars.Lbox = h.Lbox;
    pars.dim = h.dim;
    field.init(pars);
    qx_complex *p = field.psi();
    for (size_t i = 0; i < field.sites(); ++i) {
        p[i][0] = psi[2 * i];
        p[i][1] = psi[2 * i + 1];
    }
}

// src/io/io.cxx:237-273
void IO::readJaxions(Params &pars, Field &field)
{
    requireOpen();
    std::vector<double> psi;
    const ConfHeader h = readJaxionsArrays(&psi);
    pars.N = h.N;
    pars.Lbox = h.Lbox;
    pars.dim = h.dim;
    field.init(pars);
    qx_complex *p = field.psi();
    for (size_t i = 0; i < field.sites(); ++i) {
        p[i][0] = psi[2 * i];
        p[i][1] = psi[2 * i + 1];
    }
}

// src/io/io.cxx:314-370
void IO::writeConf(const Field &field, bool save_psi)
{
    requireOpen();
    const ConfHeader h{field.size(), field.dim(), field.curr(), field.time(), field.a(), field.Lbox()};
    std::vector<double> psi;
    if (save_psi) psi = psi_copy(field);
    writeConfArrays(h, field.V(), save_psi ? psi.data() : nullptr);
}

// ---- background writer -------------------------------------------------------------

SnapshotWriter::SnapshotWriter(const Field &field, bool save_psi, bool rho_from_psi)
    : save_psi_(save_psi), rho_from_psi_(rho_from_psi)
{
    const int what = ((save_psi || rho_from_psi) ? QX_SNAP_PSI : 0) | (rho_from_psi ? 0 : QX_SNAP_V);
    qx_throw(qx_snap_create(field.context(), what, &snap_));
    ctx_ = field.context();
}

SnapshotWriter::~SnapshotWriter()
{
    if (worker_.joinable()) worker_.join();
    qx_snap_destroy(snap_);
}

void SnapshotWriter::wait()
{
    if (worker_.joinable()) worker_.join();
    if (error_) {
        std::exception_ptr e = error_;
        error_ = nullptr;
        std::rethrow_exception(e);
    }
}

void SnapshotWriter::write(const Field &field, const std::string &filename)
{
    wait();
    if (field.context() != ctx_)
        throw std::invalid_argument("SnapshotWriter: write() takes the Field the writer was made for");
    field.syncToDevice();
    qx_throw(qx_snap_take(snap_));
    const ConfHeader h{field.size(), field.dim(), field.curr(), field.time(), field.a(), field.Lbox()};
    const size_t S = field.sites();
    worker_ = std::thread([this, h, S, filename]() {
        try {
            const bool want_psi = save_psi_ || rho_from_psi_;
            if (want_psi) psi_.resize(2 * S);
            if (!rho_from_psi_) V_.resize(S);
            qx_throw(qx_snap_wait(snap_, want_psi ? psi_.data() : nullptr, rho_from_psi_ ? nullptr : V_.data()));
            const double *rho = V_.data();
            if (rho_from_psi_) {
                rho_.resize(S);
                for (size_t i = 0; i < S; ++i) rho_[i] = psi_[2 * i] * psi_[2 * i] + psi_[2 * i + 1] * psi_[2 * i + 1];
                rho = rho_.data();
            }
            IO io(filename, FileMode::Create);
            io.writeConfArrays(h, rho, save_psi_ ? psi_.data() : nullptr);
            io.close();
        } catch (...) {
            error_ = std::current_exception();
        }
    });
}
