// io.h -- HDF5 configuration files (src/io/io.h:10-53, src/io/io.cxx).
// IO keeps the reference's file layout, modes, method names and
// exceptions.  The *Arrays methods are host-array forms of the same
// formats, and SnapshotWriter (no reference counterpart) writes writeConf
// files from a device snapshot on a background thread so that the time
// loop does not wait for the transfer or the disk.
//
// File layout (src/io/io.cxx:314-370): group "Header" with 1-element
// attributes N, dim, step (int) and time, a, Lbox (double); dataset
// rho/data (double, dims {N,N,N} or {N,N,1}); with save_psi, psi/real and
// psi/imag of the same shape.  Jaxions files (src/io/io.cxx:237-273):
// Header attributes N, dim (int), L (double), datasets /m and /v.
#ifndef QX_SHIM_IO_H
#define QX_SHIM_IO_H

#include <hdf5.h>

#include <exception>
#include <string>
#include <thread>
#include <vector>
#include "field.h"
#include "parse.h"

enum class FileMode { Create, ReadOnly, ReadWrite };

struct ConfHeader {
    int N = 0;
    int dim = 0;
    int step = 0;
    double time = 0.0;
    double a = 1.0;
    double Lbox = 0.0;
};

class IO {
  public:
    IO(const std::string &filename, FileMode mode = FileMode::Create);
    ~IO();
    IO(const IO &) = delete;
    IO &operator=(const IO &) = delete;

    void flush();
    void close();

    // N, dim, Lbox and psi from a writeConf file into a new Field
    void readConf(Params &p, Field &field);
    // a 2D/3D jaxions grid: psi = ai^(3/2)/sqrt(2) (m - i v), ai = 0.01
    void readJaxions(Params &p, Field &field);
    // rho/data <- V as the Field holds it; psi with save_psi
    void writeConf(const Field &field, bool save_psi = false);

    // host-array forms.  rho: S doubles; psi_c128: S interleaved complex or null
    void writeConfArrays(const ConfHeader &h, const double *rho, const double *psi_c128);
    // Header (step, time and a only if present) and, when non-null, psi and rho
    ConfHeader readConfArrays(std::vector<double> *psi_c128, std::vector<double> *rho = nullptr);
    // N, dim, L and the scaled psi of a jaxions file
    ConfHeader readJaxionsArrays(std::vector<double> *psi_c128);
    // a jaxions-layout file from m and v (test fixtures)
    void writeJaxionsArrays(int N, int dim, double L, const double *m, const double *v);

  private:
    hid_t file_ = -1;

    void requireOpen() const;
    void ensureGroup(const std::string &name);
    bool hasAttribute(const std::string &group, const std::string &name) const;
    template <typename T>
    T readAttribute(const std::string &group, const std::string &name) const;
    template <typename T>
    void writeAttribute(const std::string &group, const std::string &name, const T &value);
    void readDataset(const std::string &path, double *out, size_t count);
    void writeDataset(const std::string &group, const std::string &name, int N, int dim, const double *data);
};

class SnapshotWriter {
  public:
    // rho_from_psi: rho/data holds |psi|^2 of the snapshot instead of V
    // (for measurements that keep V, see measure()); save_psi as writeConf
    SnapshotWriter(const Field &field, bool save_psi, bool rho_from_psi = false);
    ~SnapshotWriter();
    SnapshotWriter(const SnapshotWriter &) = delete;
    SnapshotWriter &operator=(const SnapshotWriter &) = delete;

    // waits for the previous file, takes the device snapshot in stream
    // order and returns; the file is written in the background
    void write(const Field &field, const std::string &filename);
    // joins the background write and rethrows its error, if any
    void wait();
    bool rhoFromPsi() const { return rho_from_psi_; }

  private:
    qx_snap *snap_ = nullptr;
    const qx_ctx *ctx_ = nullptr;  // the Field's device state the snapshot copies
    bool save_psi_ = false;
    bool rho_from_psi_ = false;
    std::thread worker_;
    std::exception_ptr error_;
    std::vector<double> psi_, V_, rho_;
};

#endif
