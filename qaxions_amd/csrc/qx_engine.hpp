// qx_engine.hpp -- the per-(N, precision) set of FFT pass launchers.  Each
// precision is instantiated in its own translation unit (qx_lib_f32.hip,
// qx_lib_f64.hip) so the two compile in parallel.
#pragma once
#include <algorithm>
#include <memory>
#include "qx_kernels.hpp"

namespace qx {

// ---------------------------------------------------------------------------
// per-(N, precision) kernel set
// ---------------------------------------------------------------------------
struct Engine {
    virtual ~Engine() = default;
    // out: optional destination with swapped outer indices (contig_fft)
    virtual void contig(void *data, long long nlines, int op, const MulArgs &m, hipStream_t s,
                        void *out = nullptr, long long rm_m = 1, long long rm_a = 1) = 0;
    virtual void strided(void *data, long long outer, long long inner, int op, const MulArgs &m,
                         hipStream_t s) = 0;
    // drift along the strided axis (fwd, the axis' factor, inv), data columns
    // -> column tiles (see strided_fft), tile_cols() wide; inner must be a
    // multiple of it
    virtual void strided_tiles_drift(void *data, void *tiles, long long outer, long long inner,
                                     const MulArgs &m, hipStream_t s) = 0;
    virtual int tile_cols() const = 0;
    // rows of `rpm` rows per member (rpm = rows for a single lattice);
    // advance_tick: a step's density pass, which also advances the drift
    // counter (where the engine has one) in place of advance()
    virtual void r2c_rows(int mode, void *psi, const void *vin, void *vt, long long rows,
                          long long rpm, double pref, double rho_bar, hipStream_t s,
                          bool advance_tick = false) = 0;
    // C2R_SRC_TILES: psi read from `tiles`; C2R_DRIFT: the last axis' drift
    // with the per-axis table ptab (members ptab_stride apart) and 1/N
    virtual void c2r_rows(const void *vt, void *V, void *psi, long long rows, long long rpm,
                          double kick_dt, const double *kick_member, int flags,
                          unsigned long long *vmax, hipStream_t s, const void *tiles = nullptr,
                          const double2 *ptab = nullptr, const float2 *ftab = nullptr,
                          long long ptab_stride = 0, double inv_n = 0.0) = 0;
    // P(k) binned inside the last forward pass of the spectrum's transform,
    // along the contiguous lines (3D: the k1 lines of vt [k2][i0][k1]; 2D:
    // the i0 lines of vt [k1][i0]); the grid: blocks of lines, at most what
    // the GPU holds at once (the pass is a grid-stride loop)
    virtual unsigned bin_grid(int dim, unsigned cap) = 0;
    virtual void spectrum_bin(int dim, const void *vt, const BinArgs &ba, unsigned grid, hipStream_t s) = 0;
    void *tw = nullptr;  // device twiddles in stage-major order, N entries; FP32 engines then hold
                         // their N low words (see cmul2; lines under 512 points) or DITHER_K
                         // dithered tables of N entries (twd) and the drift counter (tick)
    void *twd = nullptr;
    unsigned *tick = nullptr;
    // the next drift takes the next dithered table: once after each drift
    // (all of its axes' passes), in stream order
    void advance(hipStream_t s)
    {
        if (tick) tick_advance<unsigned><<<1, 64, 0, s>>>(tick);
    }
    // streaming loads and stores in the passes (lattices whose arrays are far
    // larger than the memory-side cache; instantiated for N >= 512 only)
    bool stream_io = false;
    virtual bool has_stream_io() const = 0;
    // tile variant per strided pass kind (0 = full tile, 1 = half tile),
    // indexed by [op kind][stride class]; set by qx_tune
    int strided_variant[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    int variant(int op, const MulArgs &m, long long inner) const
    {
        const int k = op < 0 ? 0 : (op > 0 ? 1 : (m.kind == MUL_DRIFT ? 2 : 3));
        return strided_variant[k][inner > 4096 ? 1 : 0];
    }
};

template <int N, typename R> struct EngineT final : Engine {
    using C = typename cplx<R>::T;
    using G = Geom<N, R>;
    using P = Plan<N>;

    // streaming instantiations exist for N >= 512 only: no smaller N has
    // arrays large enough to gain from them
    static constexpr bool NT_OK = N >= 512;
    bool has_stream_io() const override { return NT_OK; }

    template <bool NT>
    void contig_t(void *data, long long nlines, int op, const MulArgs &m, hipStream_t s, void *out,
                  long long rm_m, long long rm_a)
    {
        const unsigned blocks = (unsigned)((nlines + G::LPB - 1) / G::LPB);
        C *d = (C *)data, *o = (C *)out;
        const C *t = (const C *)tw;
        if (op < 0) contig_fft<N, R, OP_FWD, NT><<<blocks, G::CONTIG_THREADS, 0, s>>>(d, nlines, m, t, o, rm_m, rm_a);
        else if (op > 0) contig_fft<N, R, OP_INV, NT><<<blocks, G::CONTIG_THREADS, 0, s>>>(d, nlines, m, t, o, rm_m, rm_a);
        else contig_fft<N, R, OP_POISSON, NT><<<blocks, G::CONTIG_THREADS, 0, s>>>(d, nlines, m, t, o, rm_m, rm_a);
    }
    void contig(void *data, long long nlines, int op, const MulArgs &m, hipStream_t s, void *out,
                long long rm_m, long long rm_a) override
    {
        if constexpr (NT_OK)
            if (stream_io) return contig_t<true>(data, nlines, op, m, s, out, rm_m, rm_a);
        contig_t<false>(data, nlines, op, m, s, out, rm_m, rm_a);
    }

    // writes tiles of width G::TT from blocks of G::TC columns: the tile
    // layout does not depend on tuning
    template <bool NT>
    void strided_tiles_drift_t(void *data, void *tiles, long long outer, long long inner, const MulArgs &m,
                               hipStream_t s)
    {
        constexpr int TC = G::TC, CPT = 1, TT = G::TT;
        const long long ntiles = inner / TC;
        const unsigned blocks = (unsigned)(outer * ntiles);
        strided_fft<N, R, OP_DRIFT_AXIS, TC, XM_STORE_TILES, CPT, TT, NT><<<blocks, (TC / CPT) * P::TPL, 0, s>>>(
            (C *)data, inner, ntiles, m, (const C *)tw, (C *)tiles, outer, (const C *)twd, tick);
    }
    void strided_tiles_drift(void *data, void *tiles, long long outer, long long inner, const MulArgs &m,
                             hipStream_t s) override
    {
        if constexpr (NT_OK)
            if (stream_io) return strided_tiles_drift_t<true>(data, tiles, outer, inner, m, s);
        strided_tiles_drift_t<false>(data, tiles, outer, inner, m, s);
    }
    int tile_cols() const override { return G::TT; }

    template <int TC, bool NT>
    void strided_tc(void *data, long long outer, long long inner, int op, const MulArgs &m, hipStream_t s)
    {
        const long long ntiles = (inner + TC - 1) / TC;
        const unsigned blocks = (unsigned)(outer * ntiles);
        const int threads = TC * P::TPL;
        C *d = (C *)data;
        const C *t = (const C *)tw;
        const C *td = (const C *)twd;
        if (op < 0) strided_fft<N, R, OP_FWD, TC, XM_NONE, 1, TC, NT><<<blocks, threads, 0, s>>>(d, inner, ntiles, m, t, nullptr, 0, td, tick);
        else if (op > 0) strided_fft<N, R, OP_INV, TC, XM_NONE, 1, TC, NT><<<blocks, threads, 0, s>>>(d, inner, ntiles, m, t, nullptr, 0, td, tick);
        else if (m.kind == MUL_DRIFT)
            strided_fft<N, R, OP_DRIFT_AXIS, TC, XM_NONE, 1, TC, NT><<<blocks, threads, 0, s>>>(d, inner, ntiles, m, t, nullptr, 0, td, tick);
        else strided_fft<N, R, OP_POISSON, TC, XM_NONE, 1, TC, NT><<<blocks, threads, 0, s>>>(d, inner, ntiles, m, t, nullptr, 0, td, tick);
    }
    template <bool NT>
    void strided_t(void *data, long long outer, long long inner, int op, const MulArgs &m, hipStream_t s)
    {
        if (variant(op, m, inner)) strided_tc<G::TC_HALF, NT>(data, outer, inner, op, m, s);
        else strided_tc<G::TC, NT>(data, outer, inner, op, m, s);
    }
    void strided(void *data, long long outer, long long inner, int op, const MulArgs &m,
                 hipStream_t s) override
    {
        if constexpr (NT_OK)
            if (stream_io) return strided_t<true>(data, outer, inner, op, m, s);
        strided_t<false>(data, outer, inner, op, m, s);
    }
    template <bool NT>
    void r2c_rows_t(int mode, void *psi, const void *vin, void *vt, long long rows, long long rpm,
                    double pref, double rho_bar, hipStream_t s, unsigned *tk)
    {
        const unsigned blocks = (unsigned)((rows + G::TL - 1) / G::TL);
        C *p = (C *)psi;
        const R *vi = (const R *)vin;
        C *o = (C *)vt;
        const C *t = (const C *)tw;
        if (mode == R2C_REAL)
            qx::r2c_rows<N, R, R2C_REAL, NT><<<blocks, G::ROW_THREADS, 0, s>>>(p, vi, o, rows, rpm, (R)pref, (R)rho_bar, t, tk);
        else if (mode == R2C_DENSITY)
            qx::r2c_rows<N, R, R2C_DENSITY, NT><<<blocks, G::ROW_THREADS, 0, s>>>(p, vi, o, rows, rpm, (R)pref, (R)rho_bar, t, tk);
        else
            qx::r2c_rows<N, R, R2C_DENSITY_TILES, NT><<<blocks, G::ROW_THREADS, 0, s>>>(p, vi, o, rows, rpm, (R)pref, (R)rho_bar, t, tk);
    }
    void r2c_rows(int mode, void *psi, const void *vin, void *vt, long long rows, long long rpm,
                  double pref, double rho_bar, hipStream_t s, bool advance_tick) override
    {
        unsigned *tk = advance_tick ? tick : nullptr;
        if constexpr (NT_OK)
            if (stream_io) return r2c_rows_t<true>(mode, psi, vin, vt, rows, rpm, pref, rho_bar, s, tk);
        r2c_rows_t<false>(mode, psi, vin, vt, rows, rpm, pref, rho_bar, s, tk);
    }
    long long bin_resident = 0;  // binning blocks the GPU holds at once (0: not yet queried)
    static long long bin_lines(int dim) { return dim == 3 ? (long long)(N / 2 + 1) * N : N / 2 + 1; }
    unsigned bin_grid(int dim, unsigned cap) override
    {
        // blocks beyond what the CUs hold at once would run as a late tail;
        // should the occupancy query fail, the cap stands in for it
        if (bin_resident == 0) {
            int dev = 0, cus = 0, per_cu = 0;
            const bool ok =
                hipGetDevice(&dev) == hipSuccess &&
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, contig_bin<N, R, true>, G::CONTIG_THREADS, 0) ==
                    hipSuccess &&
                cus > 0 && per_cu > 0;
            if (!ok) (void)hipGetLastError();  // not left for a later launch check to report
            bin_resident = ok ? (long long)cus * per_cu : (long long)cap;
        }
        const long long units = (bin_lines(dim) + G::LPB - 1) / G::LPB, resident = bin_resident;
        return (unsigned)std::max(1LL, std::min({units, resident, (long long)cap}));
    }
    void spectrum_bin(int dim, const void *vt, const BinArgs &ba, unsigned grid, hipStream_t s) override
    {
        if (dim == 3) contig_bin<N, R, true><<<grid, G::CONTIG_THREADS, 0, s>>>((const C *)vt, bin_lines(3), ba, (const C *)tw);
        else contig_bin<N, R, false><<<grid, G::CONTIG_THREADS, 0, s>>>((const C *)vt, bin_lines(2), ba, (const C *)tw);
    }
    void c2r_rows(const void *vt, void *V, void *psi, long long rows, long long rpm, double kick_dt,
                  const double *kick_member, int flags, unsigned long long *vmax, hipStream_t s,
                  const void *tiles, const double2 *ptab, const float2 *ftab, long long ptab_stride,
                  double inv_n) override
    {
        const unsigned blocks = (unsigned)((rows + G::TLC - 1) / G::TLC);
        auto k = qx::c2r_rows<N, R, false>;
        if constexpr (NT_OK)
            if (stream_io) k = qx::c2r_rows<N, R, true>;
        k<<<blocks, G::ROW_THREADS_C, 0, s>>>((const C *)vt, (R *)V, (C *)psi, rows, rpm, (R)kick_dt, kick_member,
                                             flags, vmax, (const C *)tw, (const C *)tiles, ptab, ftab, ptab_stride,
                                             inv_n, (const C *)twd, tick);
    }
};

template <typename R> std::unique_ptr<Engine> make_engine_t(int N)
{
    switch (N) {
    case 2: return std::make_unique<EngineT<2, R>>();
    case 4: return std::make_unique<EngineT<4, R>>();
    case 8: return std::make_unique<EngineT<8, R>>();
    case 16: return std::make_unique<EngineT<16, R>>();
    case 32: return std::make_unique<EngineT<32, R>>();
    case 64: return std::make_unique<EngineT<64, R>>();
    case 128: return std::make_unique<EngineT<128, R>>();
    case 256: return std::make_unique<EngineT<256, R>>();
    case 512: return std::make_unique<EngineT<512, R>>();
    case 1024: return std::make_unique<EngineT<1024, R>>();
    case 2048: return std::make_unique<EngineT<2048, R>>();
    case 4096: return std::make_unique<EngineT<4096, R>>();
    default: return nullptr;
    }
}

std::unique_ptr<Engine> make_engine_f32(int N);
std::unique_ptr<Engine> make_engine_f64(int N);

}  // namespace qx
