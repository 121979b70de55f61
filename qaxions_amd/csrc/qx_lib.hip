// qx_lib.hip -- host side of libqx: contexts, pass schedules, C ABI.
//
// The step schedule mirrors the reference time loop (src/main.cxx:65-85)
// and Field's operators (src/field.cxx:93-402); the fused schedule of
// qx_leapfrog merges neighbouring operators into shared HBM passes:
//
// The drift factor e^{-i dt |k|^2 / 2} is a product of one factor per axis,
// so D = D_0 D_1 D_2 with D_a = F_a^-1 e^{-i dt k_a^2 / 2} F_a / N along axis
// a alone: each axis' drift is one pass, and psi is back in real space after
// every pass.  The density then needs no inverse transform of psi.
//
//   3D step:  c2r_rows[vt -> V, kick, drift along i2]           (P end, K, D2)
//             strided drift along i1, psi rows -> column tiles psit   (D1)
//             strided drift along i0 on psit (stride N*TC)            (D0)
//             r2c_rows[psit -> density -> vt]                        (P start)
//             contig fwd i1 vt -> k-major vtk, strided Poisson on vtk,
//             contig inv vtk -> vt                                     (P)
//   2D step:  c2r_rows[kick, drift along i1] / strided drift along i0 /
//             r2c_rows[density] / contig Poisson.
// psit = [i2/TC][i0][i1][i2%TC] (TC columns per tile) keeps the i0 pass at
// stride N*TC, a few pages per block (psi's own i0 stride is N^2, one page
// per element); the next c2r_rows reads psi back from psit and writes rows.
// vt is plane-major, [i0][k2][i1], so each row block of the real<->complex
// passes touches one plane.
#include "../../include/qx.h"
#include "qx_aux.hpp"
#include "qx_engine.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using namespace qx;

namespace {

thread_local std::string g_err;

qx_status fail(qx_status s, const std::string &msg)
{
    g_err = msg;
    return s;
}

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail(QX_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_));       \
    } while (0)

constexpr double TWO_PI = 6.283185307179586476925286766559;

std::unique_ptr<Engine> make_engine(int N, bool fp64) { return fp64 ? make_engine_f64(N) : make_engine_f32(N); }

// natural index m (w^m) of each entry of the stage-major table
// (twiddle_offset); entries past the last stage stay unused
std::vector<int> twiddle_order(int N)
{
    std::vector<int> order;
    for (int s = 0; s < plan_stages(N); ++s) {
        const int R = plan_radix(N, s), Ns = plan_ns(N, s);
        if (Ns == 1) continue;
        const int tws = N / (Ns * R);
        for (int r = 1; r < R; ++r)
            for (int k = 0; k < Ns; ++k) order.push_back(r * k * tws);
    }
    return order;
}

// FP32 lines of 512 points and more: dithered drift tables (Dither)
bool dithered(int N, bool fp64) { return !fp64 && N >= 512; }

// bytes of an engine's twiddle allocation: FP64 N double2; FP32 N float2,
// then N float2 low words (lines under 512 points) or DITHER_K dithered
// tables of N float2 and the drift counter (longer lines)
size_t twiddle_bytes(int N, bool fp64)
{
    return dithered(N, fp64) ? 8 * (size_t)N * (DITHER_K + 1) + 256 : 16 * (size_t)N;
}

// a deterministic value in [-1/2, 1/2) for entry i, component j of the
// N-point tables: the starting phase of that component's sigma-delta
// sequence (splitmix64 of the three), so that the components' sequences are
// not in step with each other
double dither_phase(int N, int i, int j)
{
    uint64_t z = ((uint64_t)N << 40) ^ ((uint64_t)i << 2) ^ (uint64_t)j;
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53 - 0.5;
}

// Twiddles e^{-2 pi i m/N}, computed in FP64, in stage-major order
// (twiddle_order), uploaded to the engine's allocation (twiddle_bytes).
// FP64: the table.  FP32: the plain table, each twiddle rounded to the pair
// of neighbouring floats (cos and sin each rounded down or up) whose squared
// modulus is closest to 1 (nearest rounding of the two parts leaves
// |w|^2 - 1 with a nonzero mean over the table, -4e-9 at N = 256, a bias
// every twiddled multiply would add to the norm of the data).  Then, for
// lines under 512 points, the remainders w - wh of those rounded twiddles
// wh, rounded, which the drift's transforms add in (two-word products,
// cmul2); for longer lines the drift's DITHER_K tables: table s holds each
// component as the lower or the upper of its two neighbouring floats
// x_lo < x < x_hi, chosen by a sigma-delta sequence (acc += p with
// p = (x - x_lo)/(x_hi - x_lo), x_hi where acc reaches 1/2, which then drops
// by one), so that over any run of tables the count of x_hi follows p times
// the run's length within one; then the drift counter, zeroed.
qx_status upload_twiddles(Engine *eng, int N, bool fp64)
{
    std::vector<double> nat(2 * (size_t)N);
    for (int m = 0; m < N; ++m) {
        const double a = TWO_PI * (double)m / (double)N;
        double cr = std::cos(a), ci = -std::sin(a);
        if (4 * m == N) { cr = 0.0; ci = -1.0; }
        if (2 * m == N) { cr = -1.0; ci = 0.0; }
        if (4 * m == 3 * N) { cr = 0.0; ci = 1.0; }
        nat[2 * m] = cr;
        nat[2 * m + 1] = ci;
    }
    const std::vector<int> order = twiddle_order(N);
    if (order.size() > (size_t)N) return fail(QX_EDEVICE, "twiddle table longer than N");
    std::vector<double> tw(2 * (size_t)N, 0.0);
    for (size_t i = 0; i < order.size(); ++i) {
        tw[2 * i] = nat[2 * (size_t)order[i]];
        tw[2 * i + 1] = nat[2 * (size_t)order[i] + 1];
    }
    hipError_t e;
    if (fp64) {
        e = hipMemcpy(eng->tw, tw.data(), 16 * (size_t)N, hipMemcpyHostToDevice);
        eng->twd = nullptr;
        eng->tick = nullptr;
    } else {
        const bool dith = dithered(N, fp64);
        std::vector<float> tf(2 * (size_t)N * (dith ? DITHER_K + 1 : 2), 0.0f);
        for (int m = 0; m < N; ++m) {
            const double cr = tw[2 * m], ci = tw[2 * m + 1];
            float bc = (float)cr, bs = (float)ci;
            double best = std::fabs((double)bc * bc + (double)bs * bs - 1.0);
            const float cc[3] = {(float)cr, std::nextafterf((float)cr, -2.f), std::nextafterf((float)cr, 2.f)};
            const float ss[3] = {(float)ci, std::nextafterf((float)ci, -2.f), std::nextafterf((float)ci, 2.f)};
            for (float x : cc)
                for (float y : ss) {
                    if (std::fabs(x - cr) > 1.2e-7 || std::fabs(y - ci) > 1.2e-7) continue;  // within ~1 ulp
                    const double d = std::fabs((double)x * x + (double)y * y - 1.0);
                    if (d < best) { best = d; bc = x; bs = y; }
                }
            tf[2 * m] = bc;
            tf[2 * m + 1] = bs;
            if (!dith) {
                tf[2 * (size_t)N + 2 * m] = (float)(cr - (double)bc);
                tf[2 * (size_t)N + 2 * m + 1] = (float)(ci - (double)bs);
            }
        }
        for (size_t i = 0; dith && i < order.size(); ++i)
            for (int j = 0; j < 2; ++j) {
                const double x = tw[2 * i + j];
                float lo = (float)x;
                if ((double)lo > x) lo = std::nextafterf(lo, -2.f);
                const float hi = (double)lo == x ? lo : std::nextafterf(lo, 2.f);
                const double p = hi == lo ? 0.0 : (x - (double)lo) / ((double)hi - (double)lo);
                double acc = dither_phase(N, (int)i, j);
                for (int t = 0; t < DITHER_K; ++t) {
                    acc += p;
                    const bool up = acc >= 0.5;
                    if (up) acc -= 1.0;
                    tf[2 * (size_t)N * (1 + t) + 2 * i + j] = up ? hi : lo;
                }
            }
        e = hipMemcpy(eng->tw, tf.data(), sizeof(float) * tf.size(), hipMemcpyHostToDevice);
        eng->twd = dith ? (char *)eng->tw + 8 * (size_t)N : nullptr;
        eng->tick = dith ? (unsigned *)((char *)eng->tw + 8 * (size_t)N * (DITHER_K + 1)) : nullptr;
        if (e == hipSuccess && dith) e = hipMemset(eng->tick, 0, sizeof(unsigned));
    }
    if (e != hipSuccess) return fail(QX_EDEVICE, "twiddle upload failed");
    return QX_OK;
}

unsigned grid_for(long long n, int threads = 256, long long cap = 4096)
{
    long long b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct qx_ctx {
    int dim = 0, N = 0, device = 0;
    double L = 0.0;
    qx_precision prec = QX_FP32;
    long long S = 0, rows = 0, H1 = 0, H = 0;
    size_t csz = 0, rsz = 0;
    void *psi = nullptr, *V = nullptr, *vt = nullptr;
    void *psit = nullptr;    // 3D: psi's column tiles inside a drift (lazy)
    void *vtk = nullptr;     // 3D: k-major copy of vt inside a Poisson solve (lazy)
    void *aux_c = nullptr;   // S complex scratch (lazy)
    void *aux_vt = nullptr;  // H complex scratch (lazy)
    void *aux_r = nullptr;   // S real scratch (lazy)
    void *stage = nullptr;   // complex128 staging for host transfers (lazy)
    double *partials = nullptr;      // device reduction partials
    double *h_partials = nullptr;    // pinned host copy
    unsigned long long *vmax_bits = nullptr;
    // P(k): block histograms (up to SPEC_BLOCKS rows of N/2+1 bins) and the
    // final bins; mode counts per bin, computed once per binning (nbins,
    // spb) and kept
    double *spec_part = nullptr, *spec_P = nullptr;
    unsigned long long *spec_cnt = nullptr;
    std::vector<int64_t> spec_counts;
    int spec_key_nbins = 0, spec_key_spb = 0;
    hipStream_t stream = nullptr;
    qx_physics ph{4.0 * M_PI, 1.0, 1, 0};
    double vmax = 0.0;
    double2 *ptab = nullptr;   // drift phase table (N entries), then its FP32
                               // e^{i theta} - 1 (N float2, ftab)
    float2 *ftab = nullptr;
    double ptab_dt = std::nan("");
    // a scheduled leapfrog's per-step drift tables, SCHED_CHUNK steps at a
    // time, double-buffered: step i's N double2 then N float2 at i * N * 24
    // bytes of a buffer (lazy); pinned host staging of the same shape, and
    // per buffer the event after its last upload
    unsigned char *sched = nullptr;
    unsigned char *h_sched = nullptr;
    hipEvent_t sched_up[2] = {nullptr, nullptr};
    // captured steady-state steps for small lattices (launch-bound)
    hipGraphExec_t gexec = nullptr;
    double g_ds = std::nan("");
    qx_physics g_ph{0, 0, -1, -1};
    std::unique_ptr<Engine> eng;
};

namespace {

constexpr long long STAGE_ELEMS = 1ll << 22;  // complex128 elements per transfer chunk
constexpr int MAX_PART = 1024;
constexpr int MAX_DEVICES = 64;
constexpr int SPEC_BLOCKS = 2048;   // blocks of the binning pass (grid-stride)
constexpr long long STREAM_IO_MIN_BYTES = 512ll << 20;

qx_status ensure(qx_ctx *c, void **p, size_t bytes)
{
    if (*p) return QX_OK;
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(QX_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    (void)c;
    return QX_OK;
}

}  // namespace

// Device scratch of qx_density_stats_noalloc: the per-block (sum, max)
// partials of its reduction, 2 * MAX_PART doubles (16 KiB).  Each GPU the
// library runs on holds its own copy until the process ends.  In host code
// `scratch` names no device address: the copy of the current device is
// reached through the address hipGetSymbolAddress gives for it
// (scratch_address).
static __device__ __align__(256) double scratch[2 * MAX_PART];

namespace {

std::mutex g_scratch_mu;
double *g_scratch_addr[MAX_DEVICES] = {};

// the device address of `scratch` on `dev`, looked up once per device with
// that device current (the caller makes it current)
qx_status scratch_address(int dev, double **out)
{
    if (dev < 0 || dev >= MAX_DEVICES) return fail(QX_EINVAL, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(QX_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_addr[dev] = (double *)p;
    }
    *out = g_scratch_addr[dev];
    return QX_OK;
}

#define TRY(x)                          \
    do {                                \
        qx_status s_ = (x);             \
        if (s_ != QX_OK) return s_;     \
    } while (0)

qx_status sync(qx_ctx *c)
{
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return QX_OK;
}

bool fp64(const qx_ctx *c) { return c->prec == QX_FP64; }

// complex128 host -> device state (converted in chunks through `stage`)
qx_status upload_c(qx_ctx *c, const double *h, void *d, long long n)
{
    if (fp64(c)) {
        HIPCHK(hipMemcpyAsync(d, h, (size_t)n * 16, hipMemcpyHostToDevice, c->stream));
        return sync(c);
    }
    TRY(ensure(c, &c->stage, (size_t)STAGE_ELEMS * 16));
    for (long long off = 0; off < n; off += STAGE_ELEMS) {
        const long long m = std::min(STAGE_ELEMS, n - off);
        HIPCHK(hipMemcpyAsync(c->stage, h + 2 * off, (size_t)m * 16, hipMemcpyHostToDevice, c->stream));
        c128_to_c<float><<<grid_for(m), 256, 0, c->stream>>>((const double2 *)c->stage,
                                                            (float2 *)d + off, m);
    }
    return sync(c);
}
qx_status download_c(qx_ctx *c, const void *d, double *h, long long n)
{
    if (fp64(c)) {
        HIPCHK(hipMemcpyAsync(h, d, (size_t)n * 16, hipMemcpyDeviceToHost, c->stream));
        return sync(c);
    }
    TRY(ensure(c, &c->stage, (size_t)STAGE_ELEMS * 16));
    for (long long off = 0; off < n; off += STAGE_ELEMS) {
        const long long m = std::min(STAGE_ELEMS, n - off);
        c_to_c128<float><<<grid_for(m), 256, 0, c->stream>>>((const float2 *)d + off,
                                                            (double2 *)c->stage, m);
        HIPCHK(hipMemcpyAsync(h + 2 * off, c->stage, (size_t)m * 16, hipMemcpyDeviceToHost, c->stream));
        TRY(sync(c));
    }
    return sync(c);
}
qx_status upload_r(qx_ctx *c, const double *h, void *d, long long n)
{
    if (fp64(c)) {
        HIPCHK(hipMemcpyAsync(d, h, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
        return sync(c);
    }
    TRY(ensure(c, &c->stage, (size_t)STAGE_ELEMS * 16));
    const long long chunk = 2 * STAGE_ELEMS;
    for (long long off = 0; off < n; off += chunk) {
        const long long m = std::min(chunk, n - off);
        HIPCHK(hipMemcpyAsync(c->stage, h + off, (size_t)m * 8, hipMemcpyHostToDevice, c->stream));
        f64_to_r<float><<<grid_for(m), 256, 0, c->stream>>>((const double *)c->stage, (float *)d + off, m);
    }
    return sync(c);
}
qx_status download_r(qx_ctx *c, const void *d, double *h, long long n)
{
    if (fp64(c)) {
        HIPCHK(hipMemcpyAsync(h, d, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
        return sync(c);
    }
    TRY(ensure(c, &c->stage, (size_t)STAGE_ELEMS * 16));
    const long long chunk = 2 * STAGE_ELEMS;
    for (long long off = 0; off < n; off += chunk) {
        const long long m = std::min(chunk, n - off);
        r_to_f64<float><<<grid_for(m), 256, 0, c->stream>>>((const float *)d + off, (double *)c->stage, m);
        HIPCHK(hipMemcpyAsync(h + off, c->stage, (size_t)m * 8, hipMemcpyDeviceToHost, c->stream));
        TRY(sync(c));
    }
    return sync(c);
}

// per-axis drift phase table T[i] = e^{-i dt/2 (2 pi/L)^2 n^2}, n = fold(i),
// evaluated in FP64 on the host (libm reduces the argument exactly)
// FP32 e^{i theta} - 1 of each table entry, the difference taken in FP64
// and rounded once (as the device does when it multiplies factors in FP64)
void fill_phase_m1(float2 *F, const double2 *T, long long n)
{
    for (long long i = 0; i < n; ++i) F[i] = float2{(float)(T[i].x - 1.0), (float)T[i].y};
}

void fill_phase_table(double2 *T, int N, double L, double dt)
{
    const double dk = TWO_PI / L, c = -0.5 * dt * dk * dk;
    for (int i = 0; i < N; ++i) {
        const double n = (i <= N / 2) ? i : i - N;
        const double th = c * n * n;
        T[i] = double2{std::cos(th), std::sin(th)};
    }
}

// 3D lattices whose leading-axis drift runs on psi's column tiles
bool tiled(const qx_ctx *c) { return c->dim == 3 && c->N >= c->eng->tile_cols(); }

// the per-axis phase table for step dt (FP64 entries and their FP32
// e^{i theta} - 1), uploaded when dt changes
qx_status load_phase_table(qx_ctx *c, double dt)
{
    if (!c->ptab) {
        TRY(ensure(c, (void **)&c->ptab, (sizeof(double2) + sizeof(float2)) * c->N));
        c->ftab = (float2 *)(c->ptab + c->N);
    }
    if (!(c->ptab_dt == dt)) {
        std::vector<double2> T(c->N);
        fill_phase_table(T.data(), c->N, c->L, dt);
        std::vector<float2> F(c->N);
        fill_phase_m1(F.data(), T.data(), c->N);
        HIPCHK(hipStreamSynchronize(c->stream));  // no kernel may still read the old table
        HIPCHK(hipMemcpy(c->ptab, T.data(), sizeof(double2) * c->N, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->ftab, F.data(), sizeof(float2) * c->N, hipMemcpyHostToDevice));
        c->ptab_dt = dt;
    }
    return QX_OK;
}
// one axis' drift: factor ptab[k_a] / N on the transformed axis
MulArgs axis_drift_args(const qx_ctx *c)
{
    MulArgs m{};
    m.kind = MUL_DRIFT;
    m.geo = GEO_AXIS;
    m.ptab = c->ptab;
    m.ftab = c->ftab;
    m.inv_s = 1.0 / (double)c->N;
    m.member_div = 1;
    return m;
}
MulArgs poisson_args(const qx_ctx *c)
{
    const double dk = TWO_PI / c->L;
    MulArgs m{};
    m.kind = MUL_POISSON;
    m.geo = c->dim == 3 ? GEO_VT3 : GEO_VT2;
    m.pfac = 1.0 / (dk * dk * (double)c->S);
    m.k0fac = (c->ph.k0_mode == 1 && c->ph.rho_bar_mode == 0) ? -1.0 / (double)c->S : 0.0;
    m.member_div = c->dim == 3 ? 1 : c->H1;
    return m;
}

// ---- pass schedules ------------------------------------------------------
void c2c_passes(qx_ctx *c, void *data, int sign)
{
    MulArgs none{};
    const int N = c->N;
    c->eng->contig(data, c->rows, sign, none, c->stream);
    if (c->dim == 3) {
        c->eng->strided(data, N, N, sign, none, c->stream);
        c->eng->strided(data, 1, (long long)N * N, sign, none, c->stream);
    } else {
        c->eng->strided(data, 1, N, sign, none, c->stream);
    }
}
// the two (3D) or one (2D) leading-axis passes of psi around the drift
// D = D_2 D_1 D_0, one pass per axis (see the file header); psi in rows on
// entry and exit
qx_status drift_core(qx_ctx *c, double dt)
{
    const int N = c->N;
    TRY(load_phase_table(c, dt));
    const MulArgs am = axis_drift_args(c);
    const long long rpm = c->dim == 3 ? N : c->rows;
    c->eng->c2r_rows(nullptr, nullptr, c->psi, c->rows, rpm, 0.0, nullptr, C2R_DRIFT, nullptr, c->stream, nullptr,
                     c->ptab, c->ftab, 0, 1.0 / N);
    if (tiled(c)) {
        const long long TC = c->eng->tile_cols();
        TRY(ensure(c, &c->psit, (size_t)c->S * c->csz));
        c->eng->strided_tiles_drift(c->psi, c->psit, N, N, am, c->stream);
        c->eng->strided(c->psit, N / TC, N * TC, 0, am, c->stream);
        // back to rows: a plain copy through the tile reader
        c->eng->c2r_rows(nullptr, nullptr, c->psi, c->rows, rpm, 0.0, nullptr, C2R_SRC_TILES, nullptr, c->stream,
                         c->psit);
    } else if (c->dim == 3) {
        c->eng->strided(c->psi, N, N, 0, am, c->stream);
        c->eng->strided(c->psi, 1, (long long)N * N, 0, am, c->stream);
    } else {
        c->eng->strided(c->psi, 1, N, 0, am, c->stream);
    }
    c->eng->advance(c->stream);
    return QX_OK;
}
void vt_forward(qx_ctx *c, void *vt)
{
    MulArgs none{};
    const int N = c->N;
    if (c->dim == 3) {
        c->eng->contig(vt, c->H1 * N, -1, none, c->stream);
        c->eng->strided(vt, c->H1, N, -1, none, c->stream);
    } else {
        c->eng->contig(vt, c->H1, -1, none, c->stream);
    }
}
void vt_inverse(qx_ctx *c, void *vt)
{
    MulArgs none{};
    const int N = c->N;
    if (c->dim == 3) {
        c->eng->strided(vt, c->H1, N, +1, none, c->stream);
        c->eng->contig(vt, c->H1 * N, +1, none, c->stream);
    } else {
        c->eng->contig(vt, c->H1, +1, none, c->stream);
    }
}
// Poisson solve on the potential's half spectrum in the plane-major layout
// of c->vt (3D: [i0][k2][i1], 2D: [k1][i0]); 3D goes through the k-major
// scratch vtk = [k2][i0][k1] and back
qx_status vt_poisson(qx_ctx *c, void *vt)
{
    MulArgs none{};
    const int N = c->N;
    const MulArgs m = poisson_args(c);
    if (c->dim == 3) {
        TRY(ensure(c, &c->vtk, (size_t)c->H * c->csz));
        c->eng->contig(vt, c->H1 * N, -1, none, c->stream, c->vtk, c->H1, N);
        c->eng->strided(c->vtk, c->H1, N, 0, m, c->stream);
        c->eng->contig(c->vtk, c->H1 * N, +1, none, c->stream, vt, N, c->H1);
    } else {
        c->eng->contig(vt, c->H1, 0, m, c->stream);
    }
    return QX_OK;
}

double rho_bar_eff(const qx_ctx *c) { return c->ph.rho_bar; }

qx_status reset_vmax(qx_ctx *c)
{
    HIPCHK(hipMemsetAsync(c->vmax_bits, 0, sizeof(unsigned long long), c->stream));
    return QX_OK;
}
qx_status read_vmax(qx_ctx *c)
{
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, c->vmax_bits, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    c->vmax = v;
    return QX_OK;
}

// P pass of the reference: density, r2c, Poisson, c2r, max|V|
qx_status potential_from_psi(qx_ctx *c)
{
    TRY(reset_vmax(c));
    c->eng->r2c_rows(R2C_DENSITY, c->psi, nullptr, c->vt, c->rows, c->N, c->ph.pref, rho_bar_eff(c),
                     c->stream);
    TRY(vt_poisson(c, c->vt));
    c->eng->c2r_rows(c->vt, c->V, c->psi, c->rows, c->N, 0.0, nullptr, C2R_SRC_VT | C2R_STORE_V | C2R_VMAX,
                     c->vmax_bits, c->stream);
    return read_vmax(c);
}

template <typename R, int KIND>
qx_status reduce_t(qx_ctx *c, const void *psi, const void *V, long long n, double *sum, double *mx)
{
    const unsigned blocks = grid_for(n, RED_THREADS, MAX_PART);
    reduce_kernel<R, KIND><<<blocks, RED_THREADS, 0, c->stream>>>(
        (const typename cplx<R>::T *)psi, (const R *)V, n, c->N, c->dim, c->partials);
    HIPCHK(hipMemcpyAsync(c->h_partials, c->partials, sizeof(double) * 2 * blocks,
                          hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    double s = 0.0, m = 0.0;
    for (unsigned b = 0; b < blocks; ++b) {
        s += c->h_partials[2 * b];
        m = std::max(m, c->h_partials[2 * b + 1]);
    }
    if (sum) *sum = s;
    if (mx) *mx = m;
    return QX_OK;
}
template <int KIND>
qx_status reduce(qx_ctx *c, const void *psi, const void *V, long long n, double *sum, double *mx)
{
    return fp64(c) ? reduce_t<double, KIND>(c, psi, V, n, sum, mx)
                   : reduce_t<float, KIND>(c, psi, V, n, sum, mx);
}

bool valid_ctx(const qx_ctx *c) { return c != nullptr && c->eng != nullptr; }

// ---- the fused steady-state step, pass by pass ---------------------------
// the drift counter around work on scratch copies of the state (pass
// timing, tuning): read before, written back after, so that the steps that
// follow take the tables they would have taken without that work
qx_status tick_read(Engine *eng, hipStream_t s, unsigned *v)
{
    *v = 0;
    if (!eng->tick) return QX_OK;
    HIPCHK(hipMemcpyAsync(v, eng->tick, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return QX_OK;
}
qx_status tick_write(Engine *eng, hipStream_t s, unsigned v)
{
    if (!eng->tick) return QX_OK;
    HIPCHK(hipMemcpyAsync(eng->tick, &v, sizeof(unsigned), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return QX_OK;
}

int step_pass_count(const qx_ctx *c) { return c->dim == 3 ? 7 : 4; }

// pass p of one fused KDK step on psi (row layout), its column tiles psit,
// vt (plane-major) and vt's k-major scratch vtk; see the file header.  The
// step opens with the kick `ds` (pass 0), drifts with am's table (the
// drift_dt table of load_phase_table, or a scheduled step's own) and forms
// V with the prefactor `pref` (the density pass)
void step_pass(qx_ctx *c, int p, double ds, double pref, void *psi, void *psit, void *vt, void *vtk,
               const MulArgs &am, const MulArgs &pm)
{
    MulArgs none{};
    const int N = c->N;
    const double inv_n = 1.0 / N;
    if (c->dim == 3) {
        const bool t = tiled(c);
        const long long TC = c->eng->tile_cols();
        const int kick = C2R_SRC_VT | C2R_KICK | C2R_DRIFT | (t ? C2R_SRC_TILES : 0);
        switch (p) {
        case 0:
            c->eng->c2r_rows(vt, c->V, psi, c->rows, N, ds, nullptr, kick, nullptr, c->stream, psit, am.ptab,
                             am.ftab, 0, inv_n);
            break;
        case 1:
            if (t) c->eng->strided_tiles_drift(psi, psit, N, N, am, c->stream);
            else c->eng->strided(psi, N, N, 0, am, c->stream);
            break;
        case 2:
            if (t) c->eng->strided(psit, N / TC, N * TC, 0, am, c->stream);
            else c->eng->strided(psi, 1, (long long)N * N, 0, am, c->stream);
            break;
        case 3:   // after the step's drift: also advances the drift counter
            if (t) c->eng->r2c_rows(R2C_DENSITY_TILES, psit, nullptr, vt, c->rows, N, pref, c->ph.rho_bar, c->stream, true);
            else c->eng->r2c_rows(R2C_DENSITY, psi, nullptr, vt, c->rows, N, pref, c->ph.rho_bar, c->stream, true);
            break;
        case 4: c->eng->contig(vt, c->H1 * N, -1, none, c->stream, vtk, c->H1, N); break;
        case 5: c->eng->strided(vtk, c->H1, N, 0, pm, c->stream); break;
        case 6: c->eng->contig(vtk, c->H1 * N, +1, none, c->stream, vt, N, c->H1); break;
        }
    } else {
        const int kick = C2R_SRC_VT | C2R_KICK | C2R_DRIFT;
        switch (p) {
        case 0:
            c->eng->c2r_rows(vt, c->V, psi, c->rows, c->rows, ds, nullptr, kick, nullptr, c->stream, nullptr,
                             am.ptab, am.ftab, 0, inv_n);
            break;
        case 1: c->eng->strided(psi, 1, N, 0, am, c->stream); break;
        case 2:   // after the step's drift: also advances the drift counter
            c->eng->r2c_rows(R2C_DENSITY, psi, nullptr, vt, c->rows, c->rows, pref, c->ph.rho_bar, c->stream, true);
            break;
        case 3: c->eng->contig(vt, c->H1, 0, pm, c->stream); break;
        }
    }
}

// bytes each pass must move at minimum (every input read once, every output
// written once): psi S complex, vt H complex
double pass_bytes(const qx_ctx *c, int p)
{
    const double S = (double)c->S, H = (double)c->H, cs = (double)c->csz;
    if (c->dim == 3) {
        switch (p) {
        case 0: return H * cs + 2 * S * cs;   // read vt, rw psi
        case 1: case 2: return 2 * S * cs;
        case 3: return S * cs + H * cs;       // read psi, write vt
        default: return 2 * H * cs;
        }
    }
    switch (p) {
    case 0: return H * cs + 2 * S * cs;
    case 1: return 2 * S * cs;
    case 2: return S * cs + H * cs;
    default: return 2 * H * cs;
    }
}

#define CHECK_CTX(c)                                                  \
    do {                                                              \
        if (!valid_ctx(c)) return fail(QX_EINVAL, "null context");    \
        HIPCHK(hipSetDevice((c)->device));                            \
    } while (0)

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *qx_last_error(void) { return g_err.c_str(); }
const char *qx_version(void) { return "qx 0.1 gfx950"; }

qx_status qx_create(int dim, int N, double L, qx_precision prec, int device, qx_ctx **out)
{
    if (!out) return fail(QX_EINVAL, "out is null");
    *out = nullptr;
    if (dim != 2 && dim != 3) return fail(QX_EINVAL, "Dimension must be 2 or 3");
    if (N < 2 || (N & (N - 1)) != 0 || N > (dim == 3 ? 1024 : 4096))
        return fail(QX_EINVAL, "N must be a power of two in [2, 4096] (2D) or [2, 1024] (3D)");
    if (!(L > 0.0)) return fail(QX_EINVAL, "L must be positive");
    if (prec != QX_FP32 && prec != QX_FP64) return fail(QX_EINVAL, "unknown precision");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(QX_EINVAL, "no such device");
    HIPCHK(hipSetDevice(device));

    std::unique_ptr<qx_ctx> c(new qx_ctx());
    c->dim = dim;
    c->N = N;
    c->L = L;
    c->prec = prec;
    c->device = device;
    c->S = 1;
    for (int d = 0; d < dim; ++d) c->S *= N;
    c->rows = c->S / N;
    c->H1 = N / 2 + 1;
    c->H = c->rows * c->H1;
    c->csz = prec == QX_FP64 ? 16 : 8;
    c->rsz = prec == QX_FP64 ? 8 : 4;
    c->eng = make_engine(N, prec == QX_FP64);
    if (!c->eng) return fail(QX_EINVAL, "unsupported N");
    // streaming pass accesses for lattices whose psi is at least twice the
    // 256 MB memory-side cache (3D 512^3 and up): measured faster there,
    // slower where the arrays fit the cache (3D 256^3, 2D 4096^2)
    c->eng->stream_io = c->eng->has_stream_io() && (long long)c->S * (long long)c->csz >= STREAM_IO_MIN_BYTES;

    auto cleanup = [&](qx_status s) {
        qx_destroy(c.release());
        return s;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(QX_EDEVICE, "stream creation failed"));
    qx_status s;
    if ((s = ensure(c.get(), &c->psi, (size_t)c->S * c->csz)) != QX_OK) return cleanup(s);
    if ((s = ensure(c.get(), &c->V, (size_t)c->S * c->rsz)) != QX_OK) return cleanup(s);
    if ((s = ensure(c.get(), &c->vt, (size_t)c->H * c->csz)) != QX_OK) return cleanup(s);
    if ((s = ensure(c.get(), &c->eng->tw, twiddle_bytes(N, prec == QX_FP64))) != QX_OK) return cleanup(s);
    if ((s = ensure(c.get(), (void **)&c->partials, sizeof(double) * 2 * MAX_PART)) != QX_OK)
        return cleanup(s);
    if ((s = ensure(c.get(), (void **)&c->vmax_bits, sizeof(unsigned long long))) != QX_OK)
        return cleanup(s);
    if (hipHostMalloc((void **)&c->h_partials, sizeof(double) * 2 * MAX_PART) != hipSuccess)
        return cleanup(fail(QX_ENOMEM, "pinned host allocation failed"));
    double *scr = nullptr;  // qx_density_stats_noalloc's scratch on this device
    if ((s = scratch_address(device, &scr)) != QX_OK) return cleanup(s);

    if (upload_twiddles(c->eng.get(), N, prec == QX_FP64) != QX_OK) return cleanup(QX_EDEVICE);
    if (hipMemsetAsync(c->psi, 0, (size_t)c->S * c->csz, c->stream) != hipSuccess ||
        hipMemsetAsync(c->V, 0, (size_t)c->S * c->rsz, c->stream) != hipSuccess ||
        hipMemsetAsync(c->vt, 0, (size_t)c->H * c->csz, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return cleanup(fail(QX_EDEVICE, "state initialisation failed"));
    *out = c.release();
    return QX_OK;
}

qx_status qx_destroy(qx_ctx *c)
{
    if (!c) return QX_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void *bufs[] = {c->psi, c->V, c->vt, c->psit, c->vtk, c->aux_c, c->aux_vt, c->aux_r, c->stage,
                    c->partials, c->vmax_bits, c->spec_part, c->spec_P, c->spec_cnt, c->ptab, c->sched,
                    c->eng ? c->eng->tw : nullptr};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->h_partials) (void)hipHostFree(c->h_partials);
    if (c->h_sched) (void)hipHostFree(c->h_sched);
    for (hipEvent_t e : c->sched_up)
        if (e) (void)hipEventDestroy(e);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return QX_OK;
}

qx_status qx_set_physics(qx_ctx *c, const qx_physics *ph)
{
    if (!valid_ctx(c) || !ph) return fail(QX_EINVAL, "null argument");
    if (ph->k0_mode != 0 && ph->k0_mode != 1) return fail(QX_EINVAL, "k0_mode must be 0 or 1");
    if (ph->rho_bar_mode != 0 && ph->rho_bar_mode != 1)
        return fail(QX_EINVAL, "rho_bar_mode must be 0 or 1");
    c->ph = *ph;
    return QX_OK;
}
qx_status qx_get_physics(const qx_ctx *c, qx_physics *ph)
{
    if (!valid_ctx(c) || !ph) return fail(QX_EINVAL, "null argument");
    *ph = c->ph;
    return QX_OK;
}

qx_status qx_upload_psi(qx_ctx *c, const double *h)
{
    CHECK_CTX(c);
    if (!h) return fail(QX_EINVAL, "null psi");
    return upload_c(c, h, c->psi, c->S);
}
qx_status qx_download_psi(qx_ctx *c, double *h)
{
    CHECK_CTX(c);
    if (!h) return fail(QX_EINVAL, "null psi");
    return download_c(c, c->psi, h, c->S);
}
qx_status qx_upload_V(qx_ctx *c, const double *h)
{
    CHECK_CTX(c);
    if (!h) return fail(QX_EINVAL, "null V");
    return upload_r(c, h, c->V, c->S);
}
qx_status qx_download_V(qx_ctx *c, double *h)
{
    CHECK_CTX(c);
    if (!h) return fail(QX_EINVAL, "null V");
    return download_r(c, c->V, h, c->S);
}
qx_status qx_copy_state(qx_ctx *dst, const qx_ctx *src)
{
    if (!valid_ctx(dst) || !valid_ctx(src)) return fail(QX_EINVAL, "null context");
    if (dst->dim != src->dim || dst->N != src->N || dst->prec != src->prec)
        return fail(QX_EINVAL, "contexts differ in shape or precision");
    HIPCHK(hipSetDevice(dst->device));
    HIPCHK(hipMemcpyAsync(dst->psi, src->psi, (size_t)src->S * src->csz, hipMemcpyDeviceToDevice, dst->stream));
    HIPCHK(hipMemcpyAsync(dst->V, src->V, (size_t)src->S * src->rsz, hipMemcpyDeviceToDevice, dst->stream));
    HIPCHK(hipMemcpyAsync(dst->vt, src->vt, (size_t)src->H * src->csz, hipMemcpyDeviceToDevice, dst->stream));
    // the drift counter is part of the state: the copy's next drift takes
    // the source's next table
    if (dst->eng->tick && src->eng->tick)
        HIPCHK(hipMemcpyAsync(dst->eng->tick, src->eng->tick, sizeof(unsigned), hipMemcpyDeviceToDevice, dst->stream));
    dst->vmax = src->vmax;
    return sync(dst);
}

// ---- asynchronous snapshots ----------------------------------------------------
struct qx_snap {
    qx_ctx *ctx = nullptr;  // dereferenced only by take
    int device = 0;
    int what = 0;
    long long S = 0;
    size_t csz = 0, rsz = 0;
    void *dpsi = nullptr, *dV = nullptr;  // device copies taken on the context's stream
    void *hpsi = nullptr, *hV = nullptr;  // pinned host copies, device precision
    hipStream_t copy = nullptr;
    hipEvent_t taken = nullptr, done = nullptr;
    std::atomic<bool> pending{false};
};

qx_status qx_snap_destroy(qx_snap *s)
{
    if (!s) return QX_OK;
    (void)hipSetDevice(s->device);
    // a pending snapshot may still be copying on either stream
    if (s->pending.load() && s->taken) (void)hipEventSynchronize(s->taken);
    if (s->copy) (void)hipStreamSynchronize(s->copy);
    for (void *b : {s->dpsi, s->dV})
        if (b) (void)hipFree(b);
    for (void *b : {s->hpsi, s->hV})
        if (b) (void)hipHostFree(b);
    if (s->taken) (void)hipEventDestroy(s->taken);
    if (s->done) (void)hipEventDestroy(s->done);
    if (s->copy) (void)hipStreamDestroy(s->copy);
    delete s;
    return QX_OK;
}

qx_status qx_snap_create(qx_ctx *c, int what, qx_snap **out)
{
    if (!out) return fail(QX_EINVAL, "out is null");
    *out = nullptr;
    CHECK_CTX(c);
    if (what < 1 || what > 3) return fail(QX_EINVAL, "what: 1 psi, 2 V, 3 both");
    std::unique_ptr<qx_snap, qx_status (*)(qx_snap *)> s(new qx_snap(), qx_snap_destroy);
    s->ctx = c;
    s->device = c->device;
    s->what = what;
    s->S = c->S;
    s->csz = c->csz;
    s->rsz = c->rsz;
    if (what & QX_SNAP_PSI) {
        TRY(ensure(c, &s->dpsi, (size_t)c->S * c->csz));
        HIPCHK(hipHostMalloc(&s->hpsi, (size_t)c->S * c->csz, hipHostMallocDefault));
    }
    if (what & QX_SNAP_V) {
        TRY(ensure(c, &s->dV, (size_t)c->S * c->rsz));
        HIPCHK(hipHostMalloc(&s->hV, (size_t)c->S * c->rsz, hipHostMallocDefault));
    }
    HIPCHK(hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&s->taken, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&s->done, hipEventDisableTiming));
    *out = s.release();
    return QX_OK;
}

qx_status qx_snap_take(qx_snap *s)
{
    if (!s) return fail(QX_EINVAL, "null snapshot");
    qx_ctx *c = s->ctx;
    CHECK_CTX(c);
    if (s->pending.load()) return fail(QX_ESTATE, "the previous snapshot has not been waited for");
    // device copies in stream order: the steps queued after this call do
    // not wait for the transfer to the host
    if (s->dpsi) HIPCHK(hipMemcpyAsync(s->dpsi, c->psi, (size_t)s->S * s->csz, hipMemcpyDeviceToDevice, c->stream));
    if (s->dV) HIPCHK(hipMemcpyAsync(s->dV, c->V, (size_t)s->S * s->rsz, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipEventRecord(s->taken, c->stream));
    HIPCHK(hipStreamWaitEvent(s->copy, s->taken, 0));
    if (s->dpsi) HIPCHK(hipMemcpyAsync(s->hpsi, s->dpsi, (size_t)s->S * s->csz, hipMemcpyDeviceToHost, s->copy));
    if (s->dV) HIPCHK(hipMemcpyAsync(s->hV, s->dV, (size_t)s->S * s->rsz, hipMemcpyDeviceToHost, s->copy));
    HIPCHK(hipEventRecord(s->done, s->copy));
    s->pending.store(true);
    return QX_OK;
}

// float -> double over host threads (a 512^3 snapshot is 0.4 G values)
static void widen(const float *src, double *dst, long long n)
{
    const long long per = 1ll << 22;
    const int nt = (int)std::max(1ll, std::min<long long>(std::max(1u, std::thread::hardware_concurrency()) / 2,
                                                            std::min(16ll, (n + per - 1) / per)));
    auto part = [&](int t) {
        const long long a = n * t / nt, b = n * (t + 1) / nt;
        for (long long i = a; i < b; ++i) dst[i] = (double)src[i];
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(part, t);
    part(0);
    for (auto &th : pool) th.join();
}

qx_status qx_snap_wait(qx_snap *s, double *psi_c128, double *V)
{
    if (!s) return fail(QX_EINVAL, "null snapshot");
    if (!s->pending.load()) return fail(QX_ESTATE, "no snapshot taken");
    if ((psi_c128 && !s->hpsi) || (V && !s->hV)) return fail(QX_EINVAL, "that array was not part of the snapshot");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipEventSynchronize(s->done));
    const long long S = s->S;
    const bool f64 = s->rsz == sizeof(double);
    if (psi_c128) {
        if (f64) std::memcpy(psi_c128, s->hpsi, (size_t)S * 2 * sizeof(double));
        else widen((const float *)s->hpsi, psi_c128, 2 * S);
    }
    if (V) {
        if (f64) std::memcpy(V, s->hV, (size_t)S * sizeof(double));
        else widen((const float *)s->hV, V, S);
    }
    s->pending.store(false);
    return QX_OK;
}

qx_status qx_device_state(qx_ctx *c, void **psi, void **V)
{
    if (!valid_ctx(c)) return fail(QX_EINVAL, "null context");
    if (psi) *psi = c->psi;
    if (V) *V = c->V;
    return QX_OK;
}


// ---- split-step operators ---------------------------------------------------

qx_status qx_kick(qx_ctx *c, double dt)
{
    CHECK_CTX(c);
    if (fp64(c))
        kick_kernel<double><<<grid_for(c->S), 256, 0, c->stream>>>((double2 *)c->psi, (const double *)c->V, dt, c->S);
    else
        kick_kernel<float><<<grid_for(c->S), 256, 0, c->stream>>>((float2 *)c->psi, (const float *)c->V, (float)dt, c->S);
    return sync(c);
}

qx_status qx_drift(qx_ctx *c, double dt)
{
    CHECK_CTX(c);
    TRY(drift_core(c, dt));
    return sync(c);
}

qx_status qx_update_potential(qx_ctx *c, double *vmax)
{
    CHECK_CTX(c);
    TRY(potential_from_psi(c));
    if (vmax) *vmax = c->vmax;
    return QX_OK;
}

qx_status qx_poisson_half(qx_ctx *c, double *vh)
{
    CHECK_CTX(c);
    if (!vh) return fail(QX_EINVAL, "null half spectrum");
    TRY(ensure(c, &c->aux_vt, (size_t)c->H * c->csz));
    TRY(upload_c(c, vh, c->aux_vt, c->H));
    const MulArgs m = poisson_args(c);
    if (fp64(c))
        poisson_ref_kernel<double><<<grid_for(c->H), 256, 0, c->stream>>>((double2 *)c->aux_vt, c->rows, c->N, c->dim, m.pfac, m.k0fac);
    else
        poisson_ref_kernel<float><<<grid_for(c->H), 256, 0, c->stream>>>((float2 *)c->aux_vt, c->rows, c->N, c->dim, m.pfac, m.k0fac);
    TRY(sync(c));
    return download_c(c, c->aux_vt, vh, c->H);
}

qx_status qx_compute_energy(qx_ctx *c, double *rhomax_over_mean)
{
    CHECK_CTX(c);
    if (fp64(c))
        density_kernel<double><<<grid_for(c->S), 256, 0, c->stream>>>((const double2 *)c->psi, (double *)c->V, c->S);
    else
        density_kernel<float><<<grid_for(c->S), 256, 0, c->stream>>>((const float2 *)c->psi, (float *)c->V, c->S);
    double sum = 0.0, mx = 0.0;
    TRY(reduce<RED_VABS>(c, nullptr, c->V, c->S, &sum, &mx));
    if (rhomax_over_mean) *rhomax_over_mean = mx / (sum / (double)c->S);
    return QX_OK;
}

qx_status qx_density_stats(qx_ctx *c, double *rhomax_over_mean, double *mass)
{
    CHECK_CTX(c);
    double sum = 0.0, mx = 0.0;
    TRY(reduce<RED_RHO>(c, c->psi, nullptr, c->S, &sum, &mx));
    if (rhomax_over_mean) *rhomax_over_mean = mx / (sum / (double)c->S);
    if (mass) *mass = std::pow(c->L / c->N, c->dim) * sum;
    return QX_OK;
}

qx_status qx_density_stats_noalloc(const void *psi, int dim, int N, double L, qx_precision prec,
                                   void *stream, double *out)
{
    if (!psi || !out) return fail(QX_EINVAL, "null psi or output");
    if (dim != 2 && dim != 3) return fail(QX_EINVAL, "Dimension must be 2 or 3");
    if (N < 2 || (N & (N - 1)) != 0 || N > (dim == 3 ? 1024 : 4096))
        return fail(QX_EINVAL, "N must be a power of two in [2, 4096] (2D) or [2, 1024] (3D)");
    if (!(L > 0.0)) return fail(QX_EINVAL, "L must be positive");
    if (prec != QX_FP32 && prec != QX_FP64) return fail(QX_EINVAL, "unknown precision");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    double *scr = nullptr;
    TRY(scratch_address(dev, &scr));
    long long S = 1;
    for (int d = 0; d < dim; ++d) S *= N;
    // the grid of qx_density_stats: the same partials, summed in the same
    // order, give the same bits
    const unsigned blocks = grid_for(S, RED_THREADS, MAX_PART);
    if (blocks > (unsigned)MAX_PART) return fail(QX_EINVAL, "more partials than the scratch holds");
    hipStream_t st = (hipStream_t)stream;
    if (prec == QX_FP64)
        reduce_kernel<double, RED_RHO><<<blocks, RED_THREADS, 0, st>>>((const double2 *)psi, nullptr, S, N, dim, scr);
    else
        reduce_kernel<float, RED_RHO><<<blocks, RED_THREADS, 0, st>>>((const float2 *)psi, nullptr, S, N, dim, scr);
    density_stats_finalize<<<1, 1, 0, st>>>(scr, (int)blocks, (double)S, std::pow(L / N, dim), out);
    HIPCHK(hipGetLastError());
    return QX_OK;
}

qx_status qx_energy(qx_ctx *c, double *mass, double *kinetic, double *potential)
{
    CHECK_CTX(c);
    const double dV = std::pow(c->L / c->N, c->dim), dk = TWO_PI / c->L;
    double sum = 0.0;
    if (mass) {
        TRY(reduce<RED_RHO>(c, c->psi, nullptr, c->S, &sum, nullptr));
        *mass = dV * sum;
    }
    if (potential) {
        TRY(reduce<RED_VRHO>(c, c->psi, c->V, c->S, &sum, nullptr));
        *potential = 0.5 * dV * sum;
    }
    if (kinetic) {
        TRY(ensure(c, &c->aux_c, (size_t)c->S * c->csz));
        HIPCHK(hipMemcpyAsync(c->aux_c, c->psi, (size_t)c->S * c->csz, hipMemcpyDeviceToDevice, c->stream));
        c2c_passes(c, c->aux_c, -1);
        TRY(reduce<RED_K2>(c, c->aux_c, nullptr, c->S, &sum, nullptr));
        *kinetic = 0.5 * dV / (double)c->S * dk * dk * sum;
    }
    return QX_OK;
}

double qx_vmax(const qx_ctx *c) { return c ? c->vmax : 0.0; }

qx_status qx_spectrum(qx_ctx *c, int source, int nbins, int spb, double *Pk, int64_t *count,
                      double *kmodes)
{
    CHECK_CTX(c);
    if (source != 0 && source != 1) return fail(QX_EINVAL, "source must be 0 (V) or 1 (|psi|^2)");
    if (nbins < 1 || spb < 1 || nbins > c->N / 2 + 1)
        return fail(QX_EINVAL, "nbins must be in [1, N/2+1] and spb >= 1");
    if (!Pk || !count || !kmodes) return fail(QX_EINVAL, "null output");
    const int nsh = c->N / 2 + 1;
    TRY(ensure(c, &c->aux_vt, (size_t)c->H * c->csz));
    const unsigned grid = c->eng->bin_grid(c->dim, SPEC_BLOCKS);
    TRY(ensure(c, (void **)&c->spec_part, sizeof(double) * grid * nsh));
    TRY(ensure(c, (void **)&c->spec_P, sizeof(double) * nsh));
    TRY(ensure(c, (void **)&c->spec_cnt, sizeof(unsigned long long) * nsh));
    // the field's half spectrum: r2c along the last axis (of V, or of |psi|^2
    // formed on the fly) into [k_last][rows], then (3D) the i0 axis' strided
    // pass; the contiguous axis' transform bins the modes and writes nothing
    if (source == 0)
        c->eng->r2c_rows(R2C_REAL, nullptr, c->V, c->aux_vt, c->rows, c->rows, 1.0, 0.0, c->stream);
    else
        c->eng->r2c_rows(R2C_DENSITY, c->psi, nullptr, c->aux_vt, c->rows, c->rows, 1.0, 0.0, c->stream);
    MulArgs none{};
    if (c->dim == 3) c->eng->strided(c->aux_vt, c->H1, c->N, -1, none, c->stream);
    const BinArgs ba{nbins, spb, c->spec_part};
    c->eng->spectrum_bin(c->dim, c->aux_vt, ba, grid, c->stream);
    bin_reduce<<<nbins, 256, 0, c->stream>>>(c->spec_part, (int)grid, nbins, c->spec_P);
    HIPCHK(hipMemcpyAsync(Pk, c->spec_P, sizeof(double) * nbins, hipMemcpyDeviceToHost, c->stream));
    if (c->spec_key_nbins != nbins || c->spec_key_spb != spb) {
        // counts of this binning: once, then kept
        std::vector<unsigned long long> cnt(nbins);
        HIPCHK(hipMemsetAsync(c->spec_cnt, 0, sizeof(unsigned long long) * nbins, c->stream));
        spectrum_count_kernel<<<grid_for(c->H, 256, 2048), 256, sizeof(unsigned int) * nbins, c->stream>>>(
            c->rows, c->N, c->dim, nbins, spb, c->spec_cnt);
        HIPCHK(hipMemcpyAsync(cnt.data(), c->spec_cnt, sizeof(unsigned long long) * nbins, hipMemcpyDeviceToHost,
                              c->stream));
        TRY(sync(c));
        c->spec_counts.assign(cnt.begin(), cnt.end());
        c->spec_key_nbins = nbins;
        c->spec_key_spb = spb;
    }
    TRY(sync(c));
    const double norm = 1.0 / ((double)c->S * (double)c->S);
    const double dk = TWO_PI / c->L;
    for (int b = 0; b < nbins; ++b) {
        count[b] = c->spec_counts[b];
        Pk[b] = count[b] > 0 ? Pk[b] * norm / (double)count[b] : 0.0;
        kmodes[b] = ((double)b * spb + 0.5 * (spb - 1)) * dk;
    }
    return QX_OK;
}

// ---- step compositions --------------------------------------------------------

// n steady-state fused steps.  Lattices of at most 2^20 sites are
// launch-bound, so their steps are captured once (GRAPH_STEPS per graph,
// keyed by ds and the physics constants) and replayed.
static constexpr int GRAPH_STEPS = 8;
static constexpr long long GRAPH_MAX_SITES = 1ll << 20;

static bool same_physics(const qx_physics &a, const qx_physics &b)
{
    return a.pref == b.pref && a.rho_bar == b.rho_bar && a.k0_mode == b.k0_mode &&
           a.rho_bar_mode == b.rho_bar_mode;
}

static qx_status steady_steps(qx_ctx *c, int n, double ds, const MulArgs &am, const MulArgs &pm)
{
    const int np = step_pass_count(c);
    int done = 0;
    if (c->S <= GRAPH_MAX_SITES && n >= GRAPH_STEPS) {
        if (!c->gexec || !(c->g_ds == ds) || !same_physics(c->g_ph, c->ph)) {
            if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
            c->gexec = nullptr;
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < GRAPH_STEPS; ++i)
                for (int p = 0; p < np; ++p) step_pass(c, p, ds, c->ph.pref, c->psi, c->psit, c->vt, c->vtk, am, pm);
            HIPCHK(hipStreamEndCapture(c->stream, &g));
            const hipError_t e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(e);
            c->g_ds = ds;
            c->g_ph = c->ph;
        }
        for (; done + GRAPH_STEPS <= n; done += GRAPH_STEPS) HIPCHK(hipGraphLaunch(c->gexec, c->stream));
    }
    for (; done < n; ++done)
        for (int p = 0; p < np; ++p) step_pass(c, p, ds, c->ph.pref, c->psi, c->psit, c->vt, c->vtk, am, pm);
    return QX_OK;
}

static qx_status leapfrog_impl(qx_ctx *c, double ds, int nsteps, int first_half, int last_half, bool wait)
{
    CHECK_CTX(c);
    if (nsteps < 0) return fail(QX_EINVAL, "nsteps must be >= 0");
    const double k_first = first_half ? 0.5 * ds : ds;
    if (nsteps == 0) {
        TRY(qx_kick(c, k_first));
        if (last_half) TRY(qx_kick(c, 0.5 * ds));
        return QX_OK;
    }
    const MulArgs pm = poisson_args(c);
    TRY(load_phase_table(c, ds));
    const MulArgs am = axis_drift_args(c);
    const bool t = tiled(c);
    if (t) TRY(ensure(c, &c->psit, (size_t)c->S * c->csz));
    if (c->dim == 3) TRY(ensure(c, &c->vtk, (size_t)c->H * c->csz));
    const int np = step_pass_count(c);
    const long long rpm = c->dim == 3 ? c->N : c->rows;
    // the opening kick reads the stored V; later kicks the potential's spectrum
    c->eng->c2r_rows(nullptr, c->V, c->psi, c->rows, rpm, k_first, nullptr, C2R_KICK | C2R_DRIFT, nullptr,
                     c->stream, nullptr, c->ptab, c->ftab, 0, 1.0 / c->N);
    for (int p = 1; p < np; ++p) step_pass(c, p, ds, c->ph.pref, c->psi, c->psit, c->vt, c->vtk, am, pm);
    TRY(steady_steps(c, nsteps - 1, ds, am, pm));
    // last step's c2r: materialise V and max|V|, closing half kick; psi
    // comes back from its tiles into the row layout
    if (wait) TRY(reset_vmax(c));
    c->eng->c2r_rows(c->vt, c->V, c->psi, c->rows, rpm, 0.5 * ds, nullptr,
                     C2R_SRC_VT | C2R_STORE_V | (wait ? C2R_VMAX : 0) | (last_half ? C2R_KICK : 0) |
                         (t ? C2R_SRC_TILES : 0),
                     c->vmax_bits, c->stream, c->psit);
    return wait ? read_vmax(c) : QX_OK;
}

qx_status qx_leapfrog(qx_ctx *c, double ds, int nsteps, int first_half, int last_half)
{
    return leapfrog_impl(c, ds, nsteps, first_half, last_half, true);
}

// steps whose drift tables share one upload (and one stream wait)
static constexpr int SCHED_CHUNK = 32;

qx_status qx_leapfrog_schedule(qx_ctx *c, int nsteps, const double *drift_dt, const double *kick_dt,
                               const double *pref)
{
    CHECK_CTX(c);
    if (nsteps < 1) return fail(QX_EINVAL, "nsteps must be >= 1");
    if (!drift_dt || !kick_dt || !pref) return fail(QX_EINVAL, "null schedule array");
    for (int j = 0; j <= nsteps; ++j) {
        if (!std::isfinite(kick_dt[j])) return fail(QX_EINVAL, "kick_dt must be finite");
        if (j < nsteps && !(std::isfinite(drift_dt[j]) && std::isfinite(pref[j])))
            return fail(QX_EINVAL, "drift_dt and pref must be finite");
    }
    const int N = c->N;
    const size_t step_bytes = (sizeof(double2) + sizeof(float2)) * (size_t)N;
    const size_t buf_bytes = step_bytes * SCHED_CHUNK;
    TRY(ensure(c, (void **)&c->sched, 2 * buf_bytes));
    if (!c->h_sched) HIPCHK(hipHostMalloc((void **)&c->h_sched, 2 * buf_bytes, hipHostMallocDefault));
    for (hipEvent_t &e : c->sched_up)
        if (!e) {
            HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            HIPCHK(hipEventRecord(e, c->stream));  // a recorded state to wait on before the first upload
        }
    const bool t = tiled(c);
    if (t) TRY(ensure(c, &c->psit, (size_t)c->S * c->csz));
    if (c->dim == 3) TRY(ensure(c, &c->vtk, (size_t)c->H * c->csz));
    const MulArgs pm = poisson_args(c);
    const int np = step_pass_count(c);
    const long long rpm = c->dim == 3 ? N : c->rows;
    std::vector<double2> T(N);
    // chunk k's tables go through buffer k % 2: the host fills them while
    // the device runs chunk k - 1, and the upload is ordered after chunk
    // k - 2, the last reader of that device buffer
    for (int k = 0, j0 = 0; j0 < nsteps; ++k, j0 += SCHED_CHUNK) {
        const int m = std::min(SCHED_CHUNK, nsteps - j0), buf = k & 1;
        unsigned char *hb = c->h_sched + buf * buf_bytes, *db = c->sched + buf * buf_bytes;
        HIPCHK(hipEventSynchronize(c->sched_up[buf]));  // the upload from hb two chunks ago is done
        for (int i = 0; i < m; ++i) {
            fill_phase_table(T.data(), N, c->L, drift_dt[j0 + i]);
            unsigned char *h = hb + step_bytes * i;
            std::memcpy(h, T.data(), sizeof(double2) * N);
            fill_phase_m1((float2 *)(h + sizeof(double2) * N), T.data(), N);
        }
        HIPCHK(hipMemcpyAsync(db, hb, step_bytes * m, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipEventRecord(c->sched_up[buf], c->stream));
        for (int i = 0; i < m; ++i) {
            const int j = j0 + i;
            MulArgs am = axis_drift_args(c);
            am.ptab = (const double2 *)(db + step_bytes * i);
            am.ftab = (const float2 *)(db + step_bytes * i + sizeof(double2) * N);
            if (j == 0)  // the opening kick reads the stored V
                c->eng->c2r_rows(nullptr, c->V, c->psi, c->rows, rpm, kick_dt[0], nullptr, C2R_KICK | C2R_DRIFT,
                                 nullptr, c->stream, nullptr, am.ptab, am.ftab, 0, 1.0 / N);
            else
                step_pass(c, 0, kick_dt[j], pref[j], c->psi, c->psit, c->vt, c->vtk, am, pm);
            for (int p = 1; p < np; ++p)
                step_pass(c, p, kick_dt[j], pref[j], c->psi, c->psit, c->vt, c->vtk, am, pm);
        }
    }
    // closing c2r: V and max|V| of the last step, its kick, psi back to rows
    TRY(reset_vmax(c));
    const double kl = kick_dt[nsteps];
    c->eng->c2r_rows(c->vt, c->V, c->psi, c->rows, rpm, kl, nullptr,
                     C2R_SRC_VT | C2R_STORE_V | C2R_VMAX | (kl != 0.0 ? C2R_KICK : 0) | (t ? C2R_SRC_TILES : 0),
                     c->vmax_bits, c->stream, c->psit);
    return read_vmax(c);
}


int qx_step_pass_count(const qx_ctx *c) { return valid_ctx(c) ? step_pass_count(c) : 0; }

// which strided tile-variant slot (op kind, stride class) a step pass uses
static bool pass_slot(const qx_ctx *c, int p, int *kind, int *cls)
{
    // the class follows Engine::variant: inner stride above 4096 elements
    const long long N = c->N;
    if (c->dim == 3) {
        const bool t = tiled(c);
        switch (p) {
        case 1: *kind = 2; *cls = N > 4096 ? 1 : 0; return !t;   // drift along i1 (tile pass: fixed width)
        case 2: *kind = 2; *cls = (t ? N * c->eng->tile_cols() : N * N) > 4096 ? 1 : 0; return true;  // along i0
        case 5: *kind = 3; *cls = N > 4096 ? 1 : 0; return true;  // vt Poisson
        default: return false;
        }
    }
    if (p == 1) { *kind = 2; *cls = N > 4096 ? 1 : 0; return true; }
    return false;
}

qx_status qx_tune(qx_ctx *c, double ds, int reps, int *variants)
{
    CHECK_CTX(c);
    if (reps < 1) return fail(QX_EINVAL, "reps must be >= 1");
    const int np = step_pass_count(c);
    // two rounds with the variants alternating, the minimum per pass: a
    // single round timed the first variant while the clocks still ramped up
    double ms[2][8], t[8];
    for (int round = 0; round < 2; ++round)
        for (int v = 0; v < 2; ++v) {
            for (auto &row : c->eng->strided_variant) row[0] = row[1] = v;
            TRY(qx_time_step_passes(c, ds, reps, t, nullptr, 8));
            for (int p = 0; p < np; ++p) ms[v][p] = round ? std::min(ms[v][p], t[p]) : t[p];
        }
    for (auto &row : c->eng->strided_variant) row[0] = row[1] = 0;
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);  // captured launches used other tiles
    c->gexec = nullptr;
    for (int p = 0; p < np; ++p) {
        int k, cl;
        if (pass_slot(c, p, &k, &cl)) c->eng->strided_variant[k][cl] = ms[1][p] < ms[0][p] ? 1 : 0;
    }
    if (variants)
        for (int k = 0; k < 4; ++k)
            for (int cl = 0; cl < 2; ++cl) variants[2 * k + cl] = c->eng->strided_variant[k][cl];
    return QX_OK;
}

qx_status qx_set_stream_io(qx_ctx *c, int on, int *on_out)
{
    if (!valid_ctx(c)) return fail(QX_EINVAL, "null context");
    if (on != 0 && on != 1) return fail(QX_EINVAL, "on is 0 or 1");
    c->eng->stream_io = on && c->eng->has_stream_io();
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);  // captured launches used the other kernels
    c->gexec = nullptr;
    if (on_out) *on_out = c->eng->stream_io ? 1 : 0;
    return QX_OK;
}

qx_status qx_set_tile_variants(qx_ctx *c, const int *variants)
{
    if (!valid_ctx(c) || !variants) return fail(QX_EINVAL, "null argument");
    for (int k = 0; k < 4; ++k)
        for (int cl = 0; cl < 2; ++cl) {
            if (variants[2 * k + cl] != 0 && variants[2 * k + cl] != 1)
                return fail(QX_EINVAL, "variants are 0 (full tile) or 1 (half tile)");
            c->eng->strided_variant[k][cl] = variants[2 * k + cl];
        }
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);  // captured launches used the old tiles
    c->gexec = nullptr;
    return QX_OK;
}

const char *qx_step_pass_name(const qx_ctx *c, int p)
{
    static const char *n3[] = {"c2r_kick_drift2", "drift1_tiles", "drift0", "density_r2c",
                               "vt_fwd1", "vt_poisson0", "vt_inv1"};
    static const char *n2[] = {"c2r_kick_drift1", "drift0", "density_r2c", "vt_poisson0"};
    if (!valid_ctx(c) || p < 0 || p >= step_pass_count(c)) return "";
    return c->dim == 3 ? n3[p] : n2[p];
}

qx_status qx_time_step_passes(qx_ctx *c, double ds, int reps, double *ms, double *bytes, int cap)
{
    CHECK_CTX(c);
    const int np = step_pass_count(c);
    if (reps < 1 || !ms || cap < np) return fail(QX_EINVAL, "need reps >= 1 and room for every pass");
    unsigned tick0 = 0;
    TRY(tick_read(c->eng.get(), c->stream, &tick0));
    TRY(ensure(c, &c->aux_c, (size_t)c->S * c->csz));
    TRY(ensure(c, &c->aux_vt, (size_t)c->H * c->csz));
    HIPCHK(hipMemcpyAsync(c->aux_c, c->psi, (size_t)c->S * c->csz, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->aux_vt, c->vt, (size_t)c->H * c->csz, hipMemcpyDeviceToDevice, c->stream));
    const MulArgs pm = poisson_args(c);
    TRY(load_phase_table(c, ds));
    const MulArgs am = axis_drift_args(c);
    if (tiled(c)) TRY(ensure(c, &c->psit, (size_t)c->S * c->csz));
    if (c->dim == 3) TRY(ensure(c, &c->vtk, (size_t)c->H * c->csz));
    // pass 0 of a tiled lattice reads psi from its tiles: fill them first
    if (tiled(c)) step_pass(c, 1, ds, c->ph.pref, c->aux_c, c->psit, c->aux_vt, c->vtk, am, pm);
    // the passes inside consecutive steps, as the fused loop runs them: an
    // event before every pass and one after the last, for one warm-up step
    // and `reps` timed ones, read after a single synchronisation
    const int steps = reps + 1, nev = steps * np + 1;
    std::vector<hipEvent_t> ev(nev, nullptr);
    qx_status st = QX_OK;
    for (int i = 0; i < nev && st == QX_OK; ++i)
        if (hipEventCreate(&ev[i]) != hipSuccess) st = fail(QX_EDEVICE, "event creation failed");
    for (int k = 0; k < steps && st == QX_OK; ++k)
        for (int p = 0; p < np && st == QX_OK; ++p) {
            if (hipEventRecord(ev[k * np + p], c->stream) != hipSuccess) st = fail(QX_EDEVICE, "event record");
            step_pass(c, p, ds, c->ph.pref, c->aux_c, c->psit, c->aux_vt, c->vtk, am, pm);
        }
    if (st == QX_OK && hipEventRecord(ev[nev - 1], c->stream) != hipSuccess) st = fail(QX_EDEVICE, "event record");
    if (st == QX_OK && hipEventSynchronize(ev[nev - 1]) != hipSuccess) st = fail(QX_EDEVICE, "event timing failed");
    for (int p = 0; p < np && st == QX_OK; ++p) {
        double sum = 0.0;
        for (int k = 1; k < steps && st == QX_OK; ++k) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, ev[k * np + p], ev[k * np + p + 1]) != hipSuccess)
                st = fail(QX_EDEVICE, "event timing failed");
            sum += t;
        }
        ms[p] = sum / reps;
        if (bytes) bytes[p] = pass_bytes(c, p);
    }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (st != QX_OK) return st;
    TRY(tick_write(c->eng.get(), c->stream, tick0));
    return sync(c);
}

qx_status qx_propagate(qx_ctx *c, double ds)
{
    CHECK_CTX(c);
    TRY(qx_kick(c, 0.5 * ds));
    TRY(qx_drift(c, ds));
    TRY(potential_from_psi(c));
    return qx_kick(c, 0.5 * ds);
}

// ---- FFTBackend -------------------------------------------------------------------

qx_status qx_fft_c2c(qx_ctx *c, double *data, int sign)
{
    CHECK_CTX(c);
    if (!data) return fail(QX_EINVAL, "null data");
    if (sign != -1 && sign != 1) return fail(QX_EINVAL, "sign must be -1 or +1");
    TRY(ensure(c, &c->aux_c, (size_t)c->S * c->csz));
    TRY(upload_c(c, data, c->aux_c, c->S));
    c2c_passes(c, c->aux_c, sign);
    TRY(sync(c));
    return download_c(c, c->aux_c, data, c->S);
}

static qx_status vt_to_ref_launch(qx_ctx *c, const void *vt, void *ref)
{
    dim3 blk(32, 8), grd((unsigned)((c->rows + 31) / 32), (unsigned)((c->H1 + 31) / 32));
    if (fp64(c)) vt_to_ref<double><<<grd, blk, 0, c->stream>>>((const double2 *)vt, (double2 *)ref, c->rows, (int)c->H1);
    else vt_to_ref<float><<<grd, blk, 0, c->stream>>>((const float2 *)vt, (float2 *)ref, c->rows, (int)c->H1);
    return QX_OK;
}
static qx_status ref_to_vt_launch(qx_ctx *c, const void *ref, void *vt)
{
    dim3 blk(32, 8), grd((unsigned)((c->rows + 31) / 32), (unsigned)((c->H1 + 31) / 32));
    if (fp64(c)) ref_to_vt<double><<<grd, blk, 0, c->stream>>>((const double2 *)ref, (double2 *)vt, c->rows, (int)c->H1);
    else ref_to_vt<float><<<grd, blk, 0, c->stream>>>((const float2 *)ref, (float2 *)vt, c->rows, (int)c->H1);
    return QX_OK;
}

qx_status qx_fft_r2c(qx_ctx *c, const double *in, double *out)
{
    CHECK_CTX(c);
    if (!in || !out) return fail(QX_EINVAL, "null argument");
    TRY(ensure(c, &c->aux_r, (size_t)c->S * c->rsz));
    TRY(ensure(c, &c->aux_vt, (size_t)c->H * c->csz));
    TRY(ensure(c, &c->aux_c, (size_t)c->S * c->csz));
    TRY(upload_r(c, in, c->aux_r, c->S));
    c->eng->r2c_rows(R2C_REAL, nullptr, c->aux_r, c->aux_vt, c->rows, c->rows, 1.0, 0.0, c->stream);
    vt_forward(c, c->aux_vt);
    TRY(vt_to_ref_launch(c, c->aux_vt, c->aux_c));
    TRY(sync(c));
    return download_c(c, c->aux_c, out, c->H);
}

qx_status qx_fft_r2c_V(qx_ctx *c, double scale, double *out)
{
    CHECK_CTX(c);
    if (!out) return fail(QX_EINVAL, "null output");
    TRY(ensure(c, &c->aux_vt, (size_t)c->H * c->csz));
    TRY(ensure(c, &c->aux_c, (size_t)c->S * c->csz));
    c->eng->r2c_rows(R2C_REAL, nullptr, c->V, c->aux_vt, c->rows, c->rows, 1.0, 0.0, c->stream);
    vt_forward(c, c->aux_vt);
    TRY(vt_to_ref_launch(c, c->aux_vt, c->aux_c));
    TRY(sync(c));
    TRY(download_c(c, c->aux_c, out, c->H));
    if (scale != 1.0)
        for (long long i = 0; i < 2 * c->H; ++i) out[i] *= scale;
    return QX_OK;
}

qx_status qx_fft_c2r(qx_ctx *c, const double *in, double *out)
{
    CHECK_CTX(c);
    if (!in || !out) return fail(QX_EINVAL, "null argument");
    TRY(ensure(c, &c->aux_r, (size_t)c->S * c->rsz));
    TRY(ensure(c, &c->aux_vt, (size_t)c->H * c->csz));
    TRY(ensure(c, &c->aux_c, (size_t)c->S * c->csz));
    TRY(upload_c(c, in, c->aux_c, c->H));
    TRY(ref_to_vt_launch(c, c->aux_c, c->aux_vt));
    vt_inverse(c, c->aux_vt);
    c->eng->c2r_rows(c->aux_vt, c->aux_r, nullptr, c->rows, c->rows, 0.0, nullptr, C2R_SRC_VT | C2R_STORE_V, nullptr, c->stream);
    TRY(sync(c));
    return download_r(c, c->aux_r, out, c->S);
}

int64_t qx_sites(const qx_ctx *c) { return c ? c->S : 0; }

}  // extern "C"

namespace qx {
// stream-ordered variants without the Vmax read-back (no host wait): used by
// the corrector to chain solver steps and network passes on one stream
// the context's drift counter (device word), which picks each drift's
// dithered twiddle table; null where the lattice has no dithered tables
// (FP64, and FP32 lines under 512 points)
unsigned *drift_tick(qx_ctx *c) { return c->eng->tick; }

qx_status leapfrog_enqueue(qx_ctx *c, double ds, int nsteps, int first_half, int last_half)
{
    if (nsteps == 0) return fail(QX_EINVAL, "nsteps must be >= 1");
    return leapfrog_impl(c, ds, nsteps, first_half, last_half, false);
}
qx_status update_potential_enqueue(qx_ctx *c)
{
    CHECK_CTX(c);
    c->eng->r2c_rows(R2C_DENSITY, c->psi, nullptr, c->vt, c->rows, c->N, c->ph.pref, rho_bar_eff(c),
                     c->stream);
    TRY(vt_poisson(c, c->vt));
    c->eng->c2r_rows(c->vt, c->V, c->psi, c->rows, c->N, 0.0, nullptr, C2R_SRC_VT | C2R_STORE_V,
                     nullptr, c->stream);
    return QX_OK;
}
}  // namespace qx

extern "C" {
int64_t qx_ksites(const qx_ctx *c) { return c ? c->H : 0; }
int qx_global_N(const qx_ctx *c) { return c ? c->N : 0; }
int qx_dim(const qx_ctx *c) { return c ? c->dim : 0; }
qx_precision qx_precision_of(const qx_ctx *c) { return c ? c->prec : QX_FP32; }

// ---- initial conditions -------------------------------------------------------------

qx_status qx_ic_solitons(qx_ctx *c, const double *sol, int nsol, double bkg)
{
    CHECK_CTX(c);
    if (nsol < 0 || (nsol > 0 && !sol)) return fail(QX_EINVAL, "bad soliton table");
    TRY(ensure(c, &c->stage, (size_t)STAGE_ELEMS * 16));
    if ((size_t)nsol * 5 * sizeof(double) > (size_t)STAGE_ELEMS * 16) return fail(QX_EINVAL, "too many solitons");
    if (nsol) HIPCHK(hipMemcpyAsync(c->stage, sol, sizeof(double) * 5 * nsol, hipMemcpyHostToDevice, c->stream));
    const unsigned g = grid_for(c->S);
    if (fp64(c))
        solitons_rho_kernel<double><<<g, 256, 0, c->stream>>>((double *)c->V, (const double *)c->stage, nsol, c->N, c->dim, c->L, bkg, c->S);
    else
        solitons_rho_kernel<float><<<g, 256, 0, c->stream>>>((float *)c->V, (const double *)c->stage, nsol, c->N, c->dim, c->L, bkg, c->S);
    double sum = 0.0;
    TRY(reduce<RED_VABS>(c, nullptr, c->V, c->S, &sum, nullptr));
    const double inv_mean = (double)c->S / sum;
    if (fp64(c))
        solitons_psi_kernel<double><<<g, 256, 0, c->stream>>>((double2 *)c->psi, (const double *)c->V, inv_mean, c->S);
    else
        solitons_psi_kernel<float><<<g, 256, 0, c->stream>>>((float2 *)c->psi, (const float *)c->V, inv_mean, c->S);
    HIPCHK(hipMemsetAsync(c->V, 0, (size_t)c->S * c->rsz, c->stream));
    return sync(c);
}

qx_status qx_ic_grf(qx_ctx *c, uint64_t seed, double amp)
{
    CHECK_CTX(c);
    TRY(ensure(c, &c->aux_c, (size_t)c->S * c->csz));
    const unsigned g = grid_for(c->S);
    if (fp64(c))
        grf_noise_kernel<double><<<g, 256, 0, c->stream>>>((double2 *)c->psi, (double2 *)c->aux_c, seed, c->S);
    else
        grf_noise_kernel<float><<<g, 256, 0, c->stream>>>((float2 *)c->psi, (float2 *)c->aux_c, seed, c->S);
    c2c_passes(c, c->psi, -1);
    c2c_passes(c, c->aux_c, -1);
    if (fp64(c)) {
        radial_filter_kernel<double><<<g, 256, 0, c->stream>>>((double2 *)c->psi, c->N, c->dim, 0, c->S);
        radial_filter_kernel<double><<<g, 256, 0, c->stream>>>((double2 *)c->aux_c, c->N, c->dim, 1, c->S);
    } else {
        radial_filter_kernel<float><<<g, 256, 0, c->stream>>>((float2 *)c->psi, c->N, c->dim, 0, c->S);
        radial_filter_kernel<float><<<g, 256, 0, c->stream>>>((float2 *)c->aux_c, c->N, c->dim, 1, c->S);
    }
    c2c_passes(c, c->psi, +1);
    c2c_passes(c, c->aux_c, +1);
    const double inv_s = 1.0 / (double)c->S;
    const unsigned blocks = grid_for(c->S, RED_THREADS, MAX_PART);
    if (fp64(c))
        moments_kernel<double><<<blocks, RED_THREADS, 0, c->stream>>>((const double2 *)c->psi, inv_s, c->S, c->partials);
    else
        moments_kernel<float><<<blocks, RED_THREADS, 0, c->stream>>>((const float2 *)c->psi, inv_s, c->S, c->partials);
    HIPCHK(hipMemcpyAsync(c->h_partials, c->partials, sizeof(double) * 2 * blocks, hipMemcpyDeviceToHost, c->stream));
    TRY(sync(c));
    double s1 = 0.0, s2 = 0.0;
    for (unsigned b = 0; b < blocks; ++b) { s1 += c->h_partials[2 * b]; s2 += c->h_partials[2 * b + 1]; }
    const double mean = s1 * inv_s, var = s2 * inv_s - mean * mean;
    if (!(var > 0.0)) return fail(QX_EINVAL, "degenerate random field (zero variance)");
    if (fp64(c))
        grf_compose_kernel<double><<<g, 256, 0, c->stream>>>((double2 *)c->psi, (const double2 *)c->aux_c, inv_s, mean, 1.0 / std::sqrt(var), amp, c->S);
    else
        grf_compose_kernel<float><<<g, 256, 0, c->stream>>>((float2 *)c->psi, (const float2 *)c->aux_c, inv_s, mean, 1.0 / std::sqrt(var), amp, c->S);
    return sync(c);
}

qx_status qx_ic_blobs(qx_ctx *c, const double *tab, int nb, double bkg)
{
    CHECK_CTX(c);
    if (nb < 0 || (nb > 0 && !tab)) return fail(QX_EINVAL, "bad blob table");
    TRY(ensure(c, &c->stage, (size_t)STAGE_ELEMS * 16));
    if ((size_t)nb * 6 * sizeof(double) > (size_t)STAGE_ELEMS * 16) return fail(QX_EINVAL, "too many blobs");
    if (nb) HIPCHK(hipMemcpyAsync(c->stage, tab, sizeof(double) * 6 * nb, hipMemcpyHostToDevice, c->stream));
    const unsigned g = grid_for(c->S);
    if (fp64(c))
        blobs_kernel<double><<<g, 256, 0, c->stream>>>((double2 *)c->psi, (const double *)c->stage, nb, c->N, c->dim, c->L, bkg, c->S);
    else
        blobs_kernel<float><<<g, 256, 0, c->stream>>>((float2 *)c->psi, (const double *)c->stage, nb, c->N, c->dim, c->L, bkg, c->S);
    return sync(c);
}

// ---- plumbing -----------------------------------------------------------------------

void *qx_stream(qx_ctx *c) { return c ? (void *)c->stream : nullptr; }

qx_status qx_synchronize(qx_ctx *c)
{
    CHECK_CTX(c);
    return sync(c);
}

}  // extern "C"

// ===========================================================================
// ensembles of independent 2D lattices (no reference counterpart: the
// reference runs one lattice per process)
// ===========================================================================
namespace {

struct EnsClass {
    int N = 0, B = 0;
    long long S = 0, rows = 0, H1 = 0, H = 0;
    void *psi = nullptr, *V = nullptr, *vt = nullptr;
    double2 *d_ptab = nullptr;                                 // per member phase tables [B][N],
                                                               // then FP32 e^{i theta} - 1 [B][N]
    double *d_pfac = nullptr;                                  // per member
    double *d_kick_first = nullptr, *d_kick_mid = nullptr, *d_kick_last = nullptr;
    std::unique_ptr<Engine> eng;
    std::vector<int> members;   // ensemble member ids, slot order
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
};

float2 *ens_ftab(const EnsClass &k) { return (float2 *)(k.d_ptab + (size_t)k.B * k.N); }

}  // namespace

struct qx_ens {
    qx_precision prec = QX_FP32;
    int device = 0, n = 0;
    size_t csz = 8, rsz = 4;
    std::vector<int> N, cls, slot;
    std::vector<double> L, ds_set;
    std::vector<EnsClass> classes;
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr;
    qx_physics ph{1.0, 1.0, 1, 0};
    int graph_steps = 8;          // steps per captured graph (0: plain launches)
    hipGraphExec_t gexec = nullptr;
    int gexec_steps = 0;
    double *partials = nullptr, *h_partials = nullptr;
};

namespace {

void ens_free_graph(qx_ens *e)
{
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    e->gexec = nullptr;
    e->gexec_steps = 0;
}

// per-member constants for step size ds (array over members)
qx_status ens_set_ds(qx_ens *e, const double *ds)
{
    bool same = (int)e->ds_set.size() == e->n;
    for (int i = 0; same && i < e->n; ++i) same = e->ds_set[i] == ds[i];
    if (same) return QX_OK;
    ens_free_graph(e);
    e->ds_set.assign(ds, ds + e->n);
    for (auto &k : e->classes) {
        std::vector<double> pf(k.B), kf(k.B), km(k.B), kl(k.B);
        std::vector<double2> tab((size_t)k.B * k.N);
        for (int s = 0; s < k.B; ++s) {
            const int m = k.members[s];
            const double dk = TWO_PI / e->L[m];
            fill_phase_table(tab.data() + (size_t)s * k.N, k.N, e->L[m], ds[m]);
            pf[s] = 1.0 / (dk * dk * (double)k.S);
            km[s] = ds[m];
            kl[s] = 0.5 * ds[m];
            kf[s] = 0.5 * ds[m];
        }
        HIPCHK(hipMemcpy(k.d_ptab, tab.data(), sizeof(double2) * tab.size(), hipMemcpyHostToDevice));
        std::vector<float2> ftab(tab.size());
        fill_phase_m1(ftab.data(), tab.data(), (long long)tab.size());
        HIPCHK(hipMemcpy(ens_ftab(k), ftab.data(), sizeof(float2) * ftab.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(k.d_pfac, pf.data(), sizeof(double) * k.B, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(k.d_kick_mid, km.data(), sizeof(double) * k.B, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(k.d_kick_last, kl.data(), sizeof(double) * k.B, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(k.d_kick_first, kf.data(), sizeof(double) * k.B, hipMemcpyHostToDevice));
    }
    return QX_OK;
}

MulArgs ens_drift_args(const EnsClass &k)
{
    MulArgs m{};
    m.kind = MUL_DRIFT;
    m.geo = GEO_AXIS;        // drift along i0; i1's is in c2r_rows
    m.inv_s = 1.0 / (double)k.N;
    m.ptab = k.d_ptab;
    m.ftab = ens_ftab(k);
    m.ptab_stride = k.N;
    m.member_div = 1;
    return m;
}
MulArgs ens_poisson_args(const qx_ens *e, const EnsClass &k)
{
    MulArgs m{};
    m.kind = MUL_POISSON;
    m.geo = GEO_VT2;
    m.k0fac = (e->ph.k0_mode == 1 && e->ph.rho_bar_mode == 0) ? -1.0 / (double)k.S : 0.0;
    m.member_par = k.d_pfac;
    m.member_div = k.H1;
    return m;
}

// one fused steady-state step of one size class on stream s
void ens_class_step(qx_ens *e, EnsClass &k, hipStream_t s)
{
    const MulArgs dm = ens_drift_args(k), pm = ens_poisson_args(e, k);
    k.eng->c2r_rows(k.vt, k.V, k.psi, k.rows, k.N, 0.0, k.d_kick_mid, C2R_SRC_VT | C2R_KICK | C2R_DRIFT,
                    nullptr, s, nullptr, k.d_ptab, ens_ftab(k), k.N, 1.0 / k.N);
    k.eng->strided(k.psi, k.B, k.N, 0, dm, s);
    k.eng->r2c_rows(R2C_DENSITY, k.psi, nullptr, k.vt, k.rows, k.N, e->ph.pref, e->ph.rho_bar, s, true);
    k.eng->contig(k.vt, k.B * k.H1, 0, pm, s);
}

// `steps` steady-state steps of every class: classes run as parallel branches
qx_status ens_launch_steps(qx_ens *e, int steps)
{
    for (int i = 0; i < steps; ++i) {
        HIPCHK(hipEventRecord(e->fork, e->stream));
        for (auto &k : e->classes) {
            HIPCHK(hipStreamWaitEvent(k.stream, e->fork, 0));
            ens_class_step(e, k, k.stream);
            HIPCHK(hipEventRecord(k.done, k.stream));
        }
        for (auto &k : e->classes) HIPCHK(hipStreamWaitEvent(e->stream, k.done, 0));
    }
    return QX_OK;
}

qx_status ens_sync(qx_ens *e)
{
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    return QX_OK;
}

qx_status ens_steady_steps(qx_ens *e, int steps)
{
    if (steps <= 0) return QX_OK;
    if (e->graph_steps <= 0) return ens_launch_steps(e, steps);
    const int g = e->graph_steps;
    if (steps >= g) {
        if (!e->gexec || e->gexec_steps != g) {
            ens_free_graph(e);
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            qx_status st = ens_launch_steps(e, g);
            hipError_t ce = hipStreamEndCapture(e->stream, &graph);
            if (st != QX_OK) return st;
            if (ce != hipSuccess) return fail(QX_EDEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
            ce = hipGraphInstantiate(&e->gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ce != hipSuccess) return fail(QX_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(ce));
            e->gexec_steps = g;
        }
        for (int i = 0; i + g <= steps; i += g) HIPCHK(hipGraphLaunch(e->gexec, e->stream));
    }
    return ens_launch_steps(e, steps % g);
}

}  // namespace

extern "C" {

qx_status qx_ens_create(int n, const int *Ns, const double *Ls, qx_precision prec, int device,
                        qx_ens **out)
{
    if (!out || n < 1 || !Ns || !Ls) return fail(QX_EINVAL, "need n >= 1 members with sizes and box lengths");
    *out = nullptr;
    if (prec != QX_FP32 && prec != QX_FP64) return fail(QX_EINVAL, "unknown precision");
    for (int i = 0; i < n; ++i) {
        if (Ns[i] < 16 || (Ns[i] & (Ns[i] - 1)) || Ns[i] > 4096)
            return fail(QX_EINVAL, "ensemble members must be 2D with N a power of two in [16, 4096]");
        if (!(Ls[i] > 0.0)) return fail(QX_EINVAL, "L must be positive");
    }
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(QX_EINVAL, "no such device");
    HIPCHK(hipSetDevice(device));
    std::unique_ptr<qx_ens> e(new qx_ens());
    e->prec = prec;
    e->device = device;
    e->n = n;
    e->csz = prec == QX_FP64 ? 16 : 8;
    e->rsz = prec == QX_FP64 ? 8 : 4;
    e->N.assign(Ns, Ns + n);
    e->L.assign(Ls, Ls + n);
    e->cls.assign(n, -1);
    e->slot.assign(n, -1);
    // size classes, ordered by N (work list by size class)
    std::vector<int> sizes(Ns, Ns + n);
    std::sort(sizes.begin(), sizes.end());
    sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
    e->classes.resize(sizes.size());
    auto cleanup = [&](qx_status s) {
        qx_ens_destroy(e.release());
        return s;
    };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(QX_EDEVICE, "stream creation failed"));
    for (size_t c = 0; c < sizes.size(); ++c) {
        EnsClass &k = e->classes[c];
        k.N = sizes[c];
        for (int i = 0; i < n; ++i)
            if (Ns[i] == k.N) {
                e->cls[i] = (int)c;
                e->slot[i] = (int)k.members.size();
                k.members.push_back(i);
            }
        k.B = (int)k.members.size();
        k.S = (long long)k.N * k.N;
        k.rows = (long long)k.B * k.N;
        k.H1 = k.N / 2 + 1;
        k.H = (long long)k.N * k.H1;
        k.eng = make_engine(k.N, prec == QX_FP64);
        if (hipStreamCreateWithFlags(&k.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&k.done, hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(QX_EDEVICE, "stream creation failed"));
        void **bufs[] = {&k.psi, &k.V, &k.vt, &k.eng->tw, (void **)&k.d_ptab, (void **)&k.d_pfac,
                         (void **)&k.d_kick_first, (void **)&k.d_kick_mid, (void **)&k.d_kick_last};
        const size_t sz[] = {(size_t)(k.B * k.S) * e->csz, (size_t)(k.B * k.S) * e->rsz,
                             (size_t)(k.B * k.H) * e->csz, twiddle_bytes(k.N, prec == QX_FP64),
                             (sizeof(double2) + sizeof(float2)) * k.B * k.N, sizeof(double) * k.B, sizeof(double) * k.B,
                             sizeof(double) * k.B, sizeof(double) * k.B};
        for (int b = 0; b < 9; ++b)
            if (hipMalloc(bufs[b], sz[b]) != hipSuccess) {
                *bufs[b] = nullptr;
                return cleanup(fail(QX_ENOMEM, "ensemble allocation failed"));
            }
        if (upload_twiddles(k.eng.get(), k.N, prec == QX_FP64) != QX_OK) return cleanup(QX_EDEVICE);
        if (hipMemset(k.psi, 0, sz[0]) != hipSuccess || hipMemset(k.V, 0, sz[1]) != hipSuccess ||
            hipMemset(k.vt, 0, sz[2]) != hipSuccess)
            return cleanup(fail(QX_EDEVICE, "state initialisation failed"));
    }
    if (hipMalloc((void **)&e->partials, sizeof(double) * 2 * MAX_PART) != hipSuccess ||
        hipHostMalloc((void **)&e->h_partials, sizeof(double) * 2 * MAX_PART) != hipSuccess)
        return cleanup(fail(QX_ENOMEM, "ensemble allocation failed"));
    *out = e.release();
    return QX_OK;
}

qx_status qx_ens_destroy(qx_ens *e)
{
    if (!e) return QX_OK;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    ens_free_graph(e);
    for (auto &k : e->classes) {
        if (k.stream) (void)hipStreamSynchronize(k.stream);
        void *bufs[] = {k.psi, k.V, k.vt, k.eng ? k.eng->tw : nullptr, k.d_ptab, k.d_pfac,
                        k.d_kick_first, k.d_kick_mid, k.d_kick_last};
        for (void *b : bufs)
            if (b) (void)hipFree(b);
        if (k.done) (void)hipEventDestroy(k.done);
        if (k.stream) (void)hipStreamDestroy(k.stream);
    }
    if (e->partials) (void)hipFree(e->partials);
    if (e->h_partials) (void)hipHostFree(e->h_partials);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return QX_OK;
}

qx_status qx_ens_set_physics(qx_ens *e, const qx_physics *ph)
{
    if (!e || !ph) return fail(QX_EINVAL, "null argument");
    e->ph = *ph;
    ens_free_graph(e);
    return QX_OK;
}

qx_status qx_ens_set_graph_steps(qx_ens *e, int steps)
{
    if (!e || steps < 0) return fail(QX_EINVAL, "graph steps must be >= 0");
    if (steps != e->graph_steps) ens_free_graph(e);
    e->graph_steps = steps;
    return QX_OK;
}

static qx_status ens_member(qx_ens *e, int m, EnsClass **k, long long *off)
{
    if (!e || m < 0 || m >= e->n) return fail(QX_EINVAL, "no such member");
    HIPCHK(hipSetDevice(e->device));
    *k = &e->classes[e->cls[m]];
    *off = (long long)e->slot[m] * (*k)->S;
    return QX_OK;
}

qx_status qx_ens_upload_psi(qx_ens *e, int m, const double *psi)
{
    EnsClass *k;
    long long off;
    TRY(ens_member(e, m, &k, &off));
    if (!psi) return fail(QX_EINVAL, "null psi");
    if (e->prec == QX_FP64) {
        HIPCHK(hipMemcpy((double2 *)k->psi + off, psi, (size_t)k->S * 16, hipMemcpyHostToDevice));
        return QX_OK;
    }
    double2 *tmp = nullptr;
    HIPCHK(hipMalloc((void **)&tmp, (size_t)k->S * 16));
    hipError_t er = hipMemcpy(tmp, psi, (size_t)k->S * 16, hipMemcpyHostToDevice);
    if (er == hipSuccess) {
        c128_to_c<float><<<grid_for(k->S), 256, 0, e->stream>>>(tmp, (float2 *)k->psi + off, k->S);
        er = hipStreamSynchronize(e->stream);
    }
    (void)hipFree(tmp);
    HIPCHK(er);
    return QX_OK;
}

qx_status qx_ens_download_psi(qx_ens *e, int m, double *psi)
{
    EnsClass *k;
    long long off;
    TRY(ens_member(e, m, &k, &off));
    if (!psi) return fail(QX_EINVAL, "null psi");
    if (e->prec == QX_FP64) {
        HIPCHK(hipMemcpy(psi, (double2 *)k->psi + off, (size_t)k->S * 16, hipMemcpyDeviceToHost));
        return QX_OK;
    }
    double2 *tmp = nullptr;
    HIPCHK(hipMalloc((void **)&tmp, (size_t)k->S * 16));
    c_to_c128<float><<<grid_for(k->S), 256, 0, e->stream>>>((const float2 *)k->psi + off, tmp, k->S);
    hipError_t er = hipStreamSynchronize(e->stream);
    if (er == hipSuccess) er = hipMemcpy(psi, tmp, (size_t)k->S * 16, hipMemcpyDeviceToHost);
    (void)hipFree(tmp);
    HIPCHK(er);
    return QX_OK;
}

// each member's GRF is generated exactly as a single lattice's (qx_ic_grf)
qx_status qx_ens_ic_grf(qx_ens *e, const uint64_t *seeds, double amp)
{
    if (!e || !seeds) return fail(QX_EINVAL, "null argument");
    for (int m = 0; m < e->n; ++m) {
        EnsClass *k;
        long long off;
        TRY(ens_member(e, m, &k, &off));
        qx_ctx *c = nullptr;
        TRY(qx_create(2, k->N, e->L[m], e->prec, e->device, &c));
        qx_status st = qx_ic_grf(c, seeds[m], amp);
        if (st == QX_OK &&
            hipMemcpy((char *)k->psi + off * e->csz, c->psi, (size_t)k->S * e->csz,
                      hipMemcpyDeviceToDevice) != hipSuccess)
            st = fail(QX_EDEVICE, "member copy failed");
        qx_destroy(c);
        TRY(st);
    }
    return QX_OK;
}

// P for every member: density, r2c, Poisson, c2r (V stored)
qx_status qx_ens_update_potential(qx_ens *e)
{
    if (!e) return fail(QX_EINVAL, "null ensemble");
    HIPCHK(hipSetDevice(e->device));
    for (auto &k : e->classes) {
        const MulArgs pm = ens_poisson_args(e, k);
        if (k.d_pfac) {
            std::vector<double> pf(k.B);
            for (int s = 0; s < k.B; ++s) {
                const double dk = TWO_PI / e->L[k.members[s]];
                pf[s] = 1.0 / (dk * dk * (double)k.S);
            }
            HIPCHK(hipMemcpy(k.d_pfac, pf.data(), sizeof(double) * k.B, hipMemcpyHostToDevice));
        }
        k.eng->r2c_rows(R2C_DENSITY, k.psi, nullptr, k.vt, k.rows, k.N, e->ph.pref, e->ph.rho_bar, e->stream);
        k.eng->contig(k.vt, k.B * k.H1, 0, pm, e->stream);
        k.eng->c2r_rows(k.vt, k.V, k.psi, k.rows, k.N, 0.0, nullptr, C2R_SRC_VT | C2R_STORE_V, nullptr, e->stream);
    }
    return ens_sync(e);
}

// per size class, the faster tile variant of its strided drift pass, timed
// over `reps` steady steps of the class on saved state (restored afterwards)
qx_status qx_ens_tune(qx_ens *e, const double *ds, int reps, int *variants)
{
    if (!e || !ds) return fail(QX_EINVAL, "null argument");
    if (reps < 1) return fail(QX_EINVAL, "reps must be >= 1");
    HIPCHK(hipSetDevice(e->device));
    TRY(ens_set_ds(e, ds));
    ens_free_graph(e);   // captured launches would keep the old tiles
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    qx_status st = QX_OK;
    for (size_t ci = 0; ci < e->classes.size() && st == QX_OK; ++ci) {
        EnsClass &k = e->classes[ci];
        const size_t bp = (size_t)(k.B * k.S) * e->csz, bv = (size_t)(k.B * k.H) * e->csz;
        void *sp = nullptr, *sv = nullptr;
        unsigned tick0 = 0;
        if (st == QX_OK) st = tick_read(k.eng.get(), e->stream, &tick0);
        if (hipMalloc(&sp, bp) != hipSuccess || hipMalloc(&sv, bv) != hipSuccess) {
            st = fail(QX_ENOMEM, "tuning copies");
        } else if (hipMemcpyAsync(sp, k.psi, bp, hipMemcpyDeviceToDevice, e->stream) != hipSuccess ||
                   hipMemcpyAsync(sv, k.vt, bv, hipMemcpyDeviceToDevice, e->stream) != hipSuccess) {
            st = fail(QX_EDEVICE, "tuning copies");
        }
        float ms[2] = {0.f, 0.f};
        for (int v = 0; v < 2 && st == QX_OK; ++v) {
            for (auto &row : k.eng->strided_variant) row[0] = row[1] = v;
            ens_class_step(e, k, e->stream);   // warm
            if (hipEventRecord(e0, e->stream) != hipSuccess) st = fail(QX_EDEVICE, "event record");
            for (int r = 0; r < reps; ++r) ens_class_step(e, k, e->stream);
            if (hipEventRecord(e1, e->stream) != hipSuccess) st = fail(QX_EDEVICE, "event record");
            if (st == QX_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms[v], e0, e1) != hipSuccess))
                st = fail(QX_EDEVICE, "event timing failed");
        }
        const int best = ms[1] < ms[0] ? 1 : 0;
        for (auto &row : k.eng->strided_variant) row[0] = row[1] = best;
        if (variants) variants[ci] = best;
        if (sp && sv && st == QX_OK) {
            if (hipMemcpyAsync(k.psi, sp, bp, hipMemcpyDeviceToDevice, e->stream) != hipSuccess ||
                hipMemcpyAsync(k.vt, sv, bv, hipMemcpyDeviceToDevice, e->stream) != hipSuccess ||
                hipStreamSynchronize(e->stream) != hipSuccess)
                st = fail(QX_EDEVICE, "restoring state after tuning");
            if (st == QX_OK) st = tick_write(k.eng.get(), e->stream, tick0);
        }
        (void)hipStreamSynchronize(e->stream);
        if (sp) (void)hipFree(sp);
        if (sv) (void)hipFree(sv);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

qx_status qx_ens_set_tile_variants(qx_ens *e, const int *variants)
{
    if (!e || !variants) return fail(QX_EINVAL, "null argument");
    for (size_t ci = 0; ci < e->classes.size(); ++ci)
        if (variants[ci] != 0 && variants[ci] != 1) return fail(QX_EINVAL, "tile variant must be 0 or 1");
    ens_free_graph(e);   // captured launches would keep the old tiles
    for (size_t ci = 0; ci < e->classes.size(); ++ci)
        for (auto &row : e->classes[ci].eng->strided_variant) row[0] = row[1] = variants[ci];
    return QX_OK;
}

int qx_ens_classes(const qx_ens *e) { return e ? (int)e->classes.size() : 0; }
int qx_ens_class_size(const qx_ens *e, int ci)
{
    return (e && ci >= 0 && ci < (int)e->classes.size()) ? e->classes[ci].N : 0;
}

qx_status qx_ens_leapfrog(qx_ens *e, const double *ds, int nsteps, int first_half, int last_half)
{
    if (!e || !ds) return fail(QX_EINVAL, "null argument");
    if (nsteps < 1) return fail(QX_EINVAL, "ensemble leapfrog needs nsteps >= 1");
    HIPCHK(hipSetDevice(e->device));
    TRY(ens_set_ds(e, ds));
    // opening kick from the stored V + first step, plain launches on the main stream
    for (auto &k : e->classes) {
        std::vector<double> kf(k.B);
        for (int s = 0; s < k.B; ++s) kf[s] = (first_half ? 0.5 : 1.0) * ds[k.members[s]];
        HIPCHK(hipMemcpy(k.d_kick_first, kf.data(), sizeof(double) * k.B, hipMemcpyHostToDevice));
        const MulArgs dm = ens_drift_args(k), pm = ens_poisson_args(e, k);
        k.eng->c2r_rows(nullptr, k.V, k.psi, k.rows, k.N, 0.0, k.d_kick_first, C2R_KICK | C2R_DRIFT, nullptr,
                        e->stream, nullptr, k.d_ptab, ens_ftab(k), k.N, 1.0 / k.N);
        k.eng->strided(k.psi, k.B, k.N, 0, dm, e->stream);
        k.eng->r2c_rows(R2C_DENSITY, k.psi, nullptr, k.vt, k.rows, k.N, e->ph.pref, e->ph.rho_bar, e->stream,
                        true);
        k.eng->contig(k.vt, k.B * k.H1, 0, pm, e->stream);
    }
    TRY(ens_steady_steps(e, nsteps - 1));
    for (auto &k : e->classes)
        k.eng->c2r_rows(k.vt, k.V, k.psi, k.rows, k.N, 0.0, k.d_kick_last,
                        C2R_SRC_VT | C2R_STORE_V | (last_half ? C2R_KICK : 0), nullptr, e->stream);
    return ens_sync(e);
}

qx_status qx_ens_mass(qx_ens *e, double *mass)
{
    if (!e || !mass) return fail(QX_EINVAL, "null argument");
    for (int m = 0; m < e->n; ++m) {
        EnsClass *k;
        long long off;
        TRY(ens_member(e, m, &k, &off));
        const unsigned blocks = grid_for(k->S, RED_THREADS, MAX_PART);
        if (e->prec == QX_FP64)
            reduce_kernel<double, RED_RHO><<<blocks, RED_THREADS, 0, e->stream>>>(
                (const double2 *)k->psi + off, nullptr, k->S, k->N, 2, e->partials);
        else
            reduce_kernel<float, RED_RHO><<<blocks, RED_THREADS, 0, e->stream>>>(
                (const float2 *)k->psi + off, nullptr, k->S, k->N, 2, e->partials);
        HIPCHK(hipMemcpyAsync(e->h_partials, e->partials, sizeof(double) * 2 * blocks, hipMemcpyDeviceToHost, e->stream));
        TRY(ens_sync(e));
        double s = 0.0;
        for (unsigned b = 0; b < blocks; ++b) s += e->h_partials[2 * b];
        mass[m] = s * (e->L[m] / k->N) * (e->L[m] / k->N);
    }
    return QX_OK;
}

int qx_ens_size(const qx_ens *e) { return e ? e->n : 0; }
int64_t qx_ens_sites(const qx_ens *e)
{
    if (!e) return 0;
    int64_t s = 0;
    for (int m = 0; m < e->n; ++m) s += (int64_t)e->N[m] * e->N[m];
    return s;
}
void *qx_ens_stream(qx_ens *e) { return e ? (void *)e->stream : nullptr; }

}  // extern "C"
