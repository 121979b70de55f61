// qx_kernels.hpp -- HIP kernels of the KDK split-step path for gfx950.
//
// Layouts in HBM (d = 2 or 3, S = N^d sites, rows = S/N):
//   psi  : S complex, row-major [i0][i1][i2], last index fastest (as the
//          reference's fftw_complex array, src/field.h:92);
//   V    : S real, same order (src/field.h:94);
//   vt   : the half spectrum of the real potential, plane-major: member (2D
//          ensembles) or i0 plane (3D) outermost, then k_last, then the
//          remaining row index (vt_index); a row block of the real<->complex
//          passes then reads and writes one plane, and every other transform
//          of the Poisson solve is a contiguous or N-strided pass.  The
//          reference's own layout, [row][k_last], exists only at the API
//          boundary;
//   psit : (3D, inside a step) psi in column tiles [i2/TC][i0][i1][i2%TC]
//          (tile_index), which makes the i0 axis a column pass of stride N*TC.
//
// Pass kinds:
//   contig_fft   : lines along the contiguous axis, optionally out of place
//                  with swapped outer indices, optional fused multiply
//                  (forward, pointwise factor, inverse) for the 2D potential;
//   strided_fft  : columns at stride `inner`, optional fused one-axis drift
//                  or Poisson multiply, optionally writing column tiles;
//   r2c_rows     : density pref(|psi|^2 - rho_bar) of psi rows (or of psi's
//                  column tiles) -> vt, two real rows packed in one complex
//                  transform;
//   c2r_rows     : vt -> real rows, fused with max|V|, the kick
//                  psi <- psi e^{-i dt V} and the drift along the last axis
//                  (forward transform, factor, inverse).
#pragma once
#include "qx_fft_core.hpp"

namespace qx {

// ---------------------------------------------------------------------------
// launch geometry shared by host and device
// ---------------------------------------------------------------------------
template <int N, typename R> struct Geom {
    using P = Plan<N>;                  // column passes
    // contiguous lines (contig_fft, row kernels); the same plan (16 values per
    // thread for N >= 1024 keeps a line in one wavefront, but the row
    // kernels' registers then cost more than the block barriers it saves)
    using PR = Plan<N>;
    using C = typename cplx<R>::T;
    // contig_fft: lines per block
    static constexpr int LPB = (256 / PR::TPL) > 0 ? (256 / PR::TPL) : 1;
    static constexpr int CONTIG_THREADS = LPB * PR::TPL;
    // strided_fft: columns per block (LDS budget 64 KiB, <= 1024 threads)
    static constexpr int TC0 = (65536 / (N * (int)sizeof(C))) < 1 ? 1 : (65536 / (N * (int)sizeof(C)));
    static constexpr int TC1 = TC0 > 16 ? 16 : TC0;
    static constexpr int TC = (TC1 * P::TPL > 1024) ? (1024 / P::TPL) : TC1;
    static constexpr int STRIDED_THREADS = TC * P::TPL;
    static constexpr int TC_HALF = TC > 1 ? TC / 2 : 1;   // narrow-tile variant
    // width of psi's column tiles (3D): the strided block width, except at
    // N = 1024 in FP32, whose tiles are 16 columns wide so that the row
    // kernels read whole 128 B tile rows; its strided passes (8 columns per
    // block) then work on half tiles
    static constexpr int TT = (N == 1024 && sizeof(C) == 8) ? 16 : TC;
    // row kernels: pairs of rows per block
    static constexpr int NP0 = (256 / PR::TPL) > 8 ? (256 / PR::TPL) : 8;
    static constexpr int NP1 = (NP0 * PR::TPL > 1024) ? (1024 / PR::TPL) : NP0;
    // LDS budget of ~80 KiB: per pair max(exchange line, two half-spectrum
    // staging rows)
    static constexpr int PAIR_SLOTS = PR::PADDED > N + 2 ? PR::PADDED : N + 2;
    static constexpr int NP2 = (81920 - N * (int)sizeof(C)) / (PAIR_SLOTS * (int)sizeof(C));
    // a power of two, so row blocks tile the planes and their k-major
    // segments stay aligned
    static constexpr int NP3 = NP2 < 1 ? 1 : (NP2 < NP1 ? NP2 : NP1);
    static constexpr int NP = 1 << ilog2(NP3);
    static constexpr int TL = 2 * NP;                 // rows per block
    static constexpr int ROW_THREADS = NP * PR::TPL;
    static constexpr int XCH1 = NP * PR::PADDED;                        // one line per pair
    // staging of the half spectrum, [k][row] at pitch TL + 1 (TL >= 4): a
    // wave reading one row at 64 consecutive k then spreads over all banks
    // (at pitch TL = 16, 128 bytes, every lane hits the same bank pair)
    static constexpr int STP = TL >= 4 ? TL + 1 : TL;
    // r2c_rows: a block's segment of each k-row of vt is TL values; where
    // that is less than a 128 B line, the blocks that write the rest of the
    // line run next on the same XCD (xcd_block), so its L2 merges the parts
    static constexpr bool R2C_XCD = TL * (int)sizeof(C) < 128;
    static constexpr int STG = (N / 2 + 1) * STP;     // staging slots
    static constexpr int ROW_LDS1 = XCH1 > STG ? XCH1 : STG;    // r2c_rows

    // c2r_rows.  Where a line is one wavefront (TPL = 64) the kernel is
    // latency-limited and its waves sync line-locally: FP32 takes blocks of 8
    // waves (16 rows, whole 128 B lines of vt) with one exchange line per pair
    // (41 KB of LDS), so that 3 blocks, 6 waves per SIMD, fit a CU
    // (C2R_WAVES); 4-wave blocks of 8 rows, 6 to a CU, reading half lines of
    // vt, took 1% longer (512^3: 567-568 against 559-561 us).  FP64 keeps the
    // 4-wave blocks, as its 8-wave ones would take 82 KB.  Lines of two
    // wavefronts (N = 1024) also take 4-wave blocks, so that each exchange's
    // block barrier waits for 4 waves instead of 8, with one exchange line
    // per pair (27 KB), so that 5 blocks fit (two lines in lockstep, 45 KB,
    // left 3).  Otherwise it keeps the r2c_rows blocks and transforms a
    // pair's two psi rows in lockstep, each with its own exchange line, where
    // two lines per pair fit in LDS beside the twiddles (else one after the
    // other through one line).
    static constexpr bool WAVE_LINE = PR::TPL == 64;
    static constexpr bool PAIR_LINE = PR::TPL == 128;
    static constexpr int WAVE_PAIRS = sizeof(C) == 8 ? 8 : 4;
    static constexpr int NPC = (WAVE_LINE && NP > WAVE_PAIRS) ? WAVE_PAIRS : ((PAIR_LINE && NP > 2) ? 2 : NP);
    static constexpr int TLC = 2 * NPC;               // rows per block
    static constexpr int ROW_THREADS_C = NPC * PR::TPL;
    static constexpr bool TWO_LINES = !WAVE_LINE && !PAIR_LINE &&
        (long long)(2 * NPC * PR::PADDED + N) * (long long)sizeof(C) + 256 <= 81920;
    static constexpr int XCH = (TWO_LINES ? 2 : 1) * NPC * PR::PADDED;  // exchange slots
    static constexpr int STPC = TLC >= 4 ? TLC + 1 : TLC;
    static constexpr int STGC = (N / 2 + 1) * STPC;
    static constexpr int ROW_LDS = XCH > STGC ? XCH : STGC;
    // occupancy of the latency-limited FP32 c2r_rows: one-wavefront lines 6
    // waves per SIMD, 80 VGPRs with a few spilled (at 512^3 20 us faster than
    // 5 waves with none); two-wavefront lines 5 (90 VGPRs; 6 spilled 25 and
    // took 6.00 against 5.54 ms at 1024^3)
    static constexpr int C2R_WAVES = sizeof(C) != 8 ? 1 : (WAVE_LINE ? 6 : (PAIR_LINE ? 5 : 1));
    // one exchange line per pair: the drift transforms a pair's two rows in
    // lockstep through it (fft_line2s), so each twiddle read serves both
    // rows (1024-point lines: 5.63 -> 5.27 ms at 1024^3)
    static constexpr bool SHARED_LOCKSTEP = !TWO_LINES;
};

// pointwise factor in fused passes
enum { MUL_NONE = 0, MUL_DRIFT = 1, MUL_POISSON = 2 };
// pass operations (template argument): plain transforms, or forward,
// pointwise factor and inverse in one pass; separate instantiations keep the
// factor code out of the plain passes and name each pass in profiles
enum { OP_FWD = -1, OP_INV = 1, OP_POISSON = 2, OP_DRIFT_AXIS = 3 };
// OP_DRIFT_AXIS: one axis' drift, the factor of the line position only
// (e^{-i dt k_a^2 / 2} / N); FP32 lattices read it from the precomputed FP32
// table ftab
// geometry of the multiplied pass: how the integer frequencies of an element
// follow from (line position p, column/line index c, outer o)
enum {
    GEO_VT3 = 2,   // vt 3D, axis i0 strided inside plane o = k2, c = k1
    GEO_VT2 = 3,   // vt 2D, contiguous lines along i0, line index = k1
    GEO_AXIS = 5,  // one axis' drift: the table entry of the line position
};

struct MulArgs {
    int kind;        // MUL_*
    int geo;         // GEO_*
    // drift: per-axis FP64 phase table ptab[i] = e^{-i dt/2 (2 pi/L)^2 fold(i)^2}
    // (N entries per member, members ptab_stride apart); the factor of a mode is
    // the product over its axes, formed in FP64 and rounded once
    const double2 *ptab;
    // FP32 lattices, one axis' drift (GEO_AXIS): e^{i theta} - 1 of each
    // table entry, formed in FP64 and rounded once (same members layout)
    const float2 *ftab;
    long long ptab_stride;
    double inv_s;    // 1/S
    double pfac;     // Poisson: 1/((2 pi / L)^2 S)
    double k0fac;    // Poisson factor of the k = 0 mode
    // ensembles: per-member pfac (Poisson), member index = (outer index or
    // line index) / member_div; null for a single lattice
    const double *member_par;
    long long member_div;
};

// the multiply's constants for one member of an ensemble
__device__ __forceinline__ MulArgs member_args(const MulArgs &a, long long member)
{
    MulArgs m = a;
    if (a.kind == MUL_DRIFT) {
        m.ptab = a.ptab + member * a.ptab_stride;
        if (a.ftab) m.ftab = a.ftab + member * a.ptab_stride;
    }
    else if (a.member_par) m.pfac = a.member_par[member];
    return m;
}

// Streaming (non-temporal) stores of a pass's outputs: no pass re-reads
// what it writes, and the next pass finds it in HBM anyway at the lattice
// sizes that matter (each array is larger than L2 and the memory-side cache)
typedef float vf2_t __attribute__((ext_vector_type(2)));
typedef double vd2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_nt(float2 *p, float2 v) { __builtin_nontemporal_store(vf2_t{v.x, v.y}, (vf2_t *)p); }
__device__ __forceinline__ void st_nt(double2 *p, double2 v) { __builtin_nontemporal_store(vd2_t{v.x, v.y}, (vd2_t *)p); }
__device__ __forceinline__ void st_nt(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(double *p, double v) { __builtin_nontemporal_store(v, p); }
// Streaming loads of inputs that a pass reads once, in whole 128 B lines
// per block (an input line shared by two blocks keeps the plain load, so
// that the second block can still find it in L2)
__device__ __forceinline__ float2 ld_nt(const float2 *p)
{
    const vf2_t v = __builtin_nontemporal_load((const vf2_t *)p);
    return float2{v.x, v.y};
}
__device__ __forceinline__ double2 ld_nt(const double2 *p)
{
    const vd2_t v = __builtin_nontemporal_load((const vd2_t *)p);
    return double2{v.x, v.y};
}

// a pass's loads and stores: streaming (NT) or plain.  Streaming pays where
// the lattice's arrays are far larger than the memory-side cache; smaller
// lattices find the next pass's input there and keep plain accesses.
template <bool NT, typename T> __device__ __forceinline__ void st(T *p, T v)
{
    if constexpr (NT) st_nt(p, v);
    else *p = v;
}
template <bool NT, typename T> __device__ __forceinline__ T ld(const T *p)
{
    if constexpr (NT) return ld_nt(p);
    else return *p;
}

// index of (row = i0*N + i1, i2) of a 3D lattice in its column tiles
// [i2/TT][i0][i1][i2%TT] (TT columns per tile)
template <int N, int TT>
__device__ __forceinline__ long long tile_index(long long row, int i2)
{
    const long long i0 = row / N, i1 = row - i0 * N;
    return ((long long)(i2 / TT) * N + i0) * (long long)(N * TT) + i1 * TT + (i2 % TT);
}

// A thread's values of one row in the tiles, positions t + q*TPL: when TPL is
// a multiple of TT they are at tile_index(row, t) + q * TILE_Q, a compile-time
// stride, so the 64-bit index is formed once per row
template <typename C, int N, int TT, int TPL, bool NT, int E>
__device__ __forceinline__ void load_tile_row(C (&v)[E], const C *__restrict__ tiles, long long row, int t, bool ok)
{
    if constexpr (TPL % TT == 0) {
        const C *p = tiles + tile_index<N, TT>(row, t);
        constexpr long long QS = (long long)(TPL / TT) * N * (long long)(N * TT);
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = ok ? ld<NT>(p + q * QS) : C{0, 0};
    } else {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = ok ? ld<NT>(tiles + tile_index<N, TT>(row, t + q * TPL)) : C{0, 0};
    }
}

// index of (row, k_last) in vt for members of `rpm` rows stored one after
// another: member-major, then k-major inside a member
__device__ __forceinline__ long long vt_index(long long row, int k, long long rpm, int h1)
{
    const long long mem = row / rpm;
    return mem * (long long)h1 * rpm + (long long)k * rpm + (row - mem * rpm);
}
// the same for a block's rows r0 .. r0 + TL - 1 that lie in one member (rpm
// a multiple of TL, r0 a multiple of TL): vt_index(r0 + l, k) = base + k rpm
// + l, with the one division per block
__device__ __forceinline__ long long vt_block_base(long long r0, long long rpm, int h1)
{
    const long long mem = r0 / rpm;
    return mem * (long long)h1 * rpm + (r0 - mem * rpm);
}

__device__ __forceinline__ int foldi(int i, int N) { return (i <= N / 2) ? i : i - N; }

__device__ __forceinline__ void sincos_r(float x, float *s, float *c) { sincosf(x, s, c); }
__device__ __forceinline__ void sincos_r(double x, double *s, double *c) { sincos(x, s, c); }

// sin x and cos x - 1 of a kick angle.  |x| <= pi/4 (always so at the step
// sizes the method is run with): Taylor polynomials to x^9 / x^10, truncation
// error below 3e-9 relative; otherwise the library sincosf.  cos x - 1 is
// returned directly so that psi + psi (e^{ix} - 1) rounds once, unbiased:
// rounding cos x itself near 1 would make |e^{ix}| exceed 1 on average and
// drift the mass.
__device__ __forceinline__ void sincosm1_kick(float x, float *s, float *cm1)
{
    if (fabsf(x) <= 0.78539816f) {
        const float x2 = x * x;
        *s = x * fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, 2.7557319e-6f, -1.9841270e-4f), 8.3333333e-3f),
                               -1.6666667e-1f), 1.0f);
        *cm1 = x2 * fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, -2.7557319e-7f, 2.4801587e-5f), -1.3888889e-3f),
                                   4.1666667e-2f), -0.5f);
    } else {
        float sh, ch;
        sincosf(0.5f * x, &sh, &ch);
        *s = 2.0f * sh * ch;
        *cm1 = -2.0f * sh * sh;
    }
}
// the polynomial branch alone, for angles known to satisfy |x| <= pi/4
__device__ __forceinline__ void sincosm1_poly(float x, float *s, float *cm1)
{
    const float x2 = x * x;
    *s = x * fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, 2.7557319e-6f, -1.9841270e-4f), 8.3333333e-3f),
                           -1.6666667e-1f), 1.0f);
    *cm1 = x2 * fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, -2.7557319e-7f, 2.4801587e-5f), -1.3888889e-3f),
                               4.1666667e-2f), -0.5f);
}

// |x| <= 2^-6: sin x = x - x^3/6 and cos x - 1 = -x^2/2 + x^4/24, with
// truncation below x^4/120 <= 5e-10 and x^4/360 <= 2e-10 relative
__device__ __forceinline__ void sincosm1_short(float x, float *s, float *cm1)
{
    const float x2 = x * x;
    *s = fmaf(x * x2, -1.6666667e-1f, x);
    *cm1 = x2 * fmaf(x2, 4.1666667e-2f, -0.5f);
}

// FP64: |x| <= pi/4, Taylor polynomials to x^17 / x^18 (truncation below
// 1e-19 absolute); otherwise one sincos of x/2 gives both, sin x =
// 2 sin(x/2) cos(x/2) and cos x - 1 = -2 sin^2(x/2)
__device__ __forceinline__ void sincosm1_kick(double x, double *s, double *cm1)
{
    if (fabs(x) <= 0.78539816339744831) {
        const double y = x * x;
        double p = -1.0 / 355687428096000.0;           // -1/17!
        p = fma(p, y, 1.0 / 1307674368000.0);          //  1/15!
        p = fma(p, y, -1.0 / 6227020800.0);            // -1/13!
        p = fma(p, y, 1.0 / 39916800.0);               //  1/11!
        p = fma(p, y, -1.0 / 362880.0);                // -1/9!
        p = fma(p, y, 1.0 / 5040.0);                   //  1/7!
        p = fma(p, y, -1.0 / 120.0);                   // -1/5!
        p = fma(p, y, 1.0 / 6.0);                      //  1/3!
        *s = fma(-x * y, p, x);                        // x - x y (1/3! - y/5! + ...)
        double q = -1.0 / 6402373705728000.0;          // -1/18!
        q = fma(q, y, 1.0 / 20922789888000.0);         //  1/16!
        q = fma(q, y, -1.0 / 87178291200.0);           // -1/14!
        q = fma(q, y, 1.0 / 479001600.0);              //  1/12!
        q = fma(q, y, -1.0 / 3628800.0);               // -1/10!
        q = fma(q, y, 1.0 / 40320.0);                  //  1/8!
        q = fma(q, y, -1.0 / 720.0);                   // -1/6!
        q = fma(q, y, 1.0 / 24.0);                     //  1/4!
        q = fma(q, y, -0.5);                           // -1/2!
        *cm1 = y * q;
    } else {
        double sh, ch;
        sincos(0.5 * x, &sh, &ch);
        *s = 2.0 * sh * ch;
        *cm1 = -2.0 * sh * sh;
    }
}
// FP64: the same function (its polynomial holds to pi/4)
__device__ __forceinline__ void sincosm1_poly(double x, double *s, double *cm1) { sincosm1_kick(x, s, cm1); }
// FP64 keeps the long polynomials at every angle (the short branch is FP32's)
__device__ __forceinline__ void sincosm1_short(double x, double *s, double *cm1) { sincosm1_kick(x, s, cm1); }

// psi * e^{ix} = psi + psi (e^{ix} - 1)
template <typename C, typename R>
__device__ __forceinline__ C rotate(C a, R s, R cm1)
{
    const R tr = fma(a.x, cm1, -a.y * s), ti = fma(a.x, s, a.y * cm1);
    return C{a.x + tr, a.y + ti};
}

__device__ __forceinline__ float recip(float m) { return __builtin_amdgcn_rcpf(m); }
__device__ __forceinline__ double recip(double m) { return 1.0 / m; }

__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
    return double2{fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x)};
}

// Poisson factor -1/(k^2 S) for integer m = |n|^2 (k0fac at m = 0)
template <typename R, typename C>
__device__ __forceinline__ C poisson_factor(const MulArgs &a, int m, C x)
{
    const R f = (m == 0) ? (R)a.k0fac : -(R)a.pfac * recip((R)m);
    return C{x.x * f, x.y * f};
}

// one axis' drift factor e^{i theta} / N in FP32, from e^{i theta} - 1
// formed in FP64 and rounded once (fm1): applied as (x + x (e^{i theta} - 1))
// / N, so the result rounds once, unbiased (N is a power of two); rounding
// e^{i theta} itself near 1 would bias |psi| and drift the mass
__device__ __forceinline__ float2 drift_axis(float2 fm1, float s, float2 x)
{
    const float tr = fmaf(x.x, fm1.x, -x.y * fm1.y), ti = fmaf(x.x, fm1.y, x.y * fm1.x);
    return float2{(x.x + tr) * s, (x.y + ti) * s};
}
// FP64: the table entry tp times the column factor cf, times inv_s
__device__ __forceinline__ double2 drift_factor(const double2 &tp, const double2 &cf, double inv_s, double2 x)
{
    const double2 f = zmul(tp, cf);
    return double2{(x.x * f.x - x.y * f.y) * inv_s, (x.x * f.y + x.y * f.x) * inv_s};
}

// Work-group dispatch is round-robin over the 8 XCDs (each with its own
// L2): block b runs on XCD b % 8.  Renumber so that the blocks of one XCD
// take one contiguous run of tiles; neighbouring column tiles then share
// their L2 (a 128 B line split across tiles is fetched once).
__device__ __forceinline__ long long xcd_block()
{
    const long long nb = gridDim.x, b = blockIdx.x;
    if (nb % 8 != 0) return b;
    return (b % 8) * (nb / 8) + b / 8;
}

template <int N, typename C>
__device__ __forceinline__ void load_tw(C *tw, const C *twg)
{
    for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = twg[i];
}
// FP32 twiddles with their low words (twg[N + m], see cmul2), interleaved
// for TW_HI / TW_TWO transforms: tw[2m], tw[2m + 1]
__device__ __forceinline__ void load_tw2(float2 *tw, const float2 *twg, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const float2 h = twg[i], l = twg[n + i];
        reinterpret_cast<float4 *>(tw)[i] = float4{h.x, h.y, l.x, l.y};
    }
}

// FP32 drift transforms of lines of 512 points and more (one-word
// twiddles, see TwoWord) take their twiddles from one of DITHER_K tables,
// table (tick mod DITHER_K) for the drift applied when the context's drift
// counter reads `tick`.  Each twiddle component runs through a sigma-delta
// sequence of its two neighbouring floats over the tables, so its running
// mean is the exact value and the sum of its errors over any run of steps
// stays within about one unit in the last place.  One rounded table is the
// same operator error in every step: F^-1 e^{i theta} F applied to psi step
// after step then drifts psi's norm (mass) and phase linearly in the step
// count; over the dithered tables these errors stay bounded, at the cost of
// one-word products.  The host builds the tables and advances the counter
// once per drift (tick_advance).
constexpr int DITHER_K = 256;
template <typename C>
__device__ __forceinline__ const C *drift_table(const C *twd, const unsigned *tick, int n)
{
    return twd + (long long)(*tick & (DITHER_K - 1)) * n;
}
template <typename U>  // unsigned (a template: one definition across translation units)
__global__ void tick_advance(U *tick)
{
    if (threadIdx.x == 0) *tick += 1u;
}

// ---------------------------------------------------------------------------
// contiguous lines: data[line*N + pos], LPB lines per block.  With `out`, the
// result of line l = a*rm_m + b (b < rm_m) goes to out line b*rm_a + a: the
// two outer indices of a 3-index array swap places (rm_a = count of a), so a
// block's LPB lines land in LPB whole lines of the other layout
// ---------------------------------------------------------------------------
template <int N, typename R, int OP, bool NT>  // OP_FWD, OP_INV or OP_POISSON (2D vt lines)
__global__ void __launch_bounds__((Geom<N, R>::CONTIG_THREADS))
contig_fft(typename cplx<R>::T *__restrict__ data, long long nlines, MulArgs mul,
           const typename cplx<R>::T *__restrict__ twg, typename cplx<R>::T *__restrict__ out,
           long long rm_m, long long rm_a)
{
    using C = typename cplx<R>::T;
    using P = typename Geom<N, R>::PR;
    using G = Geom<N, R>;
    __shared__ C lds[G::LPB * P::PADDED];
    __shared__ C tw[N];
    load_tw<N>(tw, twg);
    const int t = threadIdx.x % P::TPL, l = threadIdx.x / P::TPL;
    const long long line = (long long)blockIdx.x * G::LPB + l;
    const bool ok = line < nlines;
    C *mine = lds + l * P::PADDED;
    C v[P::E];
    C *src = data + line * N;
#pragma unroll
    for (int q = 0; q < P::E; ++q) v[q] = ok ? ld<NT>(src + t + q * P::TPL) : C{0, 0};
    __syncthreads();
    if constexpr (OP == OP_FWD) {
        fft_line<N, -1, LdsContig>(v, t, mine, 1, tw);
    } else if constexpr (OP == OP_INV) {
        fft_line<N, +1, LdsContig>(v, t, mine, 1, tw);
    } else {
        static_assert(OP == OP_POISSON, "contiguous passes: fwd, inv or the 2D Poisson solve");
        fft_line<N, -1, LdsContig>(v, t, mine, 1, tw);
        // GEO_VT2: position = k0 (folded), line = member * (N/2+1) + k1
        const int k1 = (int)(line % mul.member_div);
        const MulArgs ma = member_args(mul, line / mul.member_div);
#pragma unroll
        for (int q = 0; q < P::E; ++q) {
            const int n0 = foldi(t + q * P::TPL, N);
            v[q] = poisson_factor<R>(ma, n0 * n0 + k1 * k1, v[q]);
        }
        fft_line<N, +1, LdsContig>(v, t, mine, 1, tw);
    }
    if (ok) {
        C *dst = out ? out + ((line % rm_m) * rm_a + line / rm_m) * N : src;
#pragma unroll
        for (int q = 0; q < P::E; ++q) st<NT>(dst + t + q * P::TPL, v[q]);
    }
}

// ---------------------------------------------------------------------------
// strided columns: data[o*N*inner + p*inner + col], col in [0, inner); a
// block takes TC adjacent columns of one outer index.  TC is a tile variant
// (Geom::TC or half of it): narrower tiles use half the LDS, so twice as
// many blocks share a CU and overlap their loads with each other's
// transforms; the variant is picked per pass by measurement.
//
// Out-of-place tile mode (3D psi): a block's N x TC tile of columns
// (col0..col0+TC-1 of outer index o) is written contiguously to `tiles` at
// ((col0/TC)*outer + o)*N*TC, as [pos][col % TC].  With data =
// psi[i0][i1][i2] (o = i0, col = i2, the pass along i1) the tiles form
// [i2/TC][i0][i1][i2%TC] (see tile_index): their i0 axis is a column pass of
// stride N*TC, 128-byte segments in 16 pages per block, where psi's own i0
// axis (stride N^2) would take one page per element.
// ---------------------------------------------------------------------------
enum { XM_NONE = 0, XM_STORE_TILES = 1 };

// CPT columns per thread: a block of TC/CPT thread columns transforms its
// CPT columns one after another through one LDS area (CPT = 2 keeps a
// 16-column tile's 128-byte row segments with 512 threads per block).
// NT: streaming accesses where a block row is a whole 128 B line (narrower
// rows share lines between blocks, which then merge in L2)
// Two-word twiddles (cmul2) in the FP32 drift transforms of lines shorter
// than 512 points.  Lines of 512 points and more take dithered one-word
// tables instead (Dither): the second word's products cost 2.5% of the 512^3
// step (2.30 against 2.36 ms, same box; the drift passes 442 against 476 us)
// and 11% of the 1024^3 step (23.6 -> 26.3 ms), while the dithered tables
// keep psi's error against FP64 where the two words put it (512^3, 50
// steps: 1.3766e-6 against 1.3764e-6).  The mass drift is larger (3.8e-9
// against 1.7e-10 over those 50 steps) but bounded: each twiddle's error
// sums to about one unit in the last place over the tables.
template <int N, typename R> struct TwoWord {
    static constexpr bool value = std::is_same<R, float>::value && N < 512;
};
// FP32 drift transforms of the longer lines: dithered one-word tables
// (drift_table)
template <int N, typename R> struct Dither {
    static constexpr bool value = std::is_same<R, float>::value && N >= 512;
};

// Waves per SIMD that a strided FP32 drift block's LDS allows: its two-word
// twiddles raise the register count, and the hint keeps it within what that
// occupancy leaves (1 = no hint for every other pass)
template <int N, typename R, int OP, int TC, int CPT> struct StridedWaves {
    static constexpr bool ON = OP == OP_DRIFT_AXIS && TwoWord<N, R>::value;
    // exchange area and twiddles; the FP32 drift table where it fits
    static constexpr long long LDS0 = 8ll * N * (TC / CPT) + (ON ? 16ll : 8ll) * N;
    static constexpr bool FTAB_LDS = LDS0 + 8ll * N <= 81920;
    static constexpr long long LDS = LDS0 + (FTAB_LDS ? 8ll * N : 0);
    static constexpr int W0 = (int)(163840 / LDS) * ((TC / CPT) * Plan<N>::TPL / 64) / 4;
    static constexpr int value = !ON ? 1 : (W0 < 1 ? 1 : (W0 > 8 ? 8 : W0));
};

template <int N, typename R, int OP, int TC, int XM = XM_NONE, int CPT = 1, int TT = TC, bool NT = false>  // OP_FWD, OP_INV, OP_DRIFT_AXIS or OP_POISSON
__global__ void __launch_bounds__(((TC / CPT) * Plan<N>::TPL))
__attribute__((amdgpu_waves_per_eu(StridedWaves<N, R, OP, TC, CPT>::value)))
strided_fft(typename cplx<R>::T *__restrict__ data, long long inner, long long ntiles,
            MulArgs mul, const typename cplx<R>::T *__restrict__ twg,
            typename cplx<R>::T *__restrict__ tiles, long long outer,
            const typename cplx<R>::T *__restrict__ twd, const unsigned *__restrict__ tick)
{
    using C = typename cplx<R>::T;
    using P = Plan<N>;
    constexpr int TW = TC / CPT;
    // FP32 drift: the transforms act on psi every step and take two-word
    // twiddles (TW_TWO, interleaved in one table) where TwoWord holds
    constexpr bool FAX = OP == OP_DRIFT_AXIS && std::is_same<R, float>::value;
    constexpr bool XTW = FAX && TwoWord<N, R>::value;
    constexpr int TWM = XTW ? TW_TWO : TW_PLAIN;
    __shared__ C lds[N * TW];
    __shared__ __align__(16) C tw[XTW ? 2 * N : N];
    // the drift's phase table is staged in LDS when small: an FP32
    // single-axis drift stages its FP32 table where two blocks still fit a
    // CU (otherwise each thread reads its entries from global memory, where
    // they stay cached; loading them ahead into registers would cost the
    // second block)
    constexpr bool PT_LDS = OP == OP_DRIFT_AXIS && (XTW ? StridedWaves<N, R, OP, TC, CPT>::FTAB_LDS : N <= 1024);
    __shared__ double2 ptab_lds[PT_LDS ? (FAX ? N / 2 : N) : 1];
    if constexpr (XTW) load_tw2((float2 *)tw, (const float2 *)twg, N);
    else if constexpr (FAX && Dither<N, R>::value) load_tw<N>(tw, drift_table(twd, tick, N));
    else load_tw<N>(tw, twg);
    const int c = threadIdx.x % TW, t = threadIdx.x / TW;
    const long long b = xcd_block();
    const long long o = b / ntiles;
    const long long col0 = (b % ntiles) * TC;
    const MulArgs ma = member_args(mul, o / (mul.member_div > 0 ? mul.member_div : 1));
    const double2 *ptab = ma.ptab;
    const float2 *ftab = ma.ftab;
    if constexpr (PT_LDS && FAX) {
        float2 *fl = (float2 *)ptab_lds;
        for (int i = threadIdx.x; i < N; i += blockDim.x) fl[i] = ma.ftab[i];
        ftab = fl;
    } else if constexpr (PT_LDS) {
        for (int i = threadIdx.x; i < N; i += blockDim.x) ptab_lds[i] = ma.ptab[i];
        ptab = ptab_lds;
    }
    C *const tile0 = tiles + ((col0 / TT) * outer + o) * (long long)N * TT + (col0 % TT) + c;
    C *const base0 = data + o * (long long)N * inner + col0 + c;
    C v[CPT][P::E];
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const bool ok = col0 + c + j * TW < inner;
#pragma unroll
        for (int q = 0; q < P::E; ++q) {
            const C *src = base0 + (long long)(t + q * P::TPL) * inner + j * TW;
            v[j][q] = ok ? ld<NT && TC * sizeof(C) >= 128>(src) : C{0, 0};
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        const long long col = col0 + c + j * TW;
        const bool ok = col < inner;
        if constexpr (OP == OP_FWD) {
            fft_line<N, -1, LdsCols>(v[j], t, lds + c, TW, tw);
        } else if constexpr (OP == OP_INV) {
            fft_line<N, +1, LdsCols>(v[j], t, lds + c, TW, tw);
        } else {
            fft_line<N, -1, LdsCols, TWM>(v[j], t, lds + c, TW, tw);
            const int cc = ok ? (int)col : 0;
            if constexpr (FAX) {
                const float s = (float)ma.inv_s;
#pragma unroll
                for (int q = 0; q < P::E; ++q) v[j][q] = drift_axis(ftab[t + q * P::TPL], s, v[j][q]);
            } else if constexpr (OP == OP_DRIFT_AXIS) {
                const double2 one{1.0, 0.0};
#pragma unroll
                for (int q = 0; q < P::E; ++q) v[j][q] = drift_factor(ptab[t + q * P::TPL], one, ma.inv_s, v[j][q]);
            } else {  // GEO_VT3 Poisson
                const int a = foldi(cc, N);
                const int mc = a * a + (int)(o * o);
#pragma unroll
                for (int q = 0; q < P::E; ++q) {
                    const int n0 = foldi(t + q * P::TPL, N);
                    v[j][q] = poisson_factor<R>(ma, n0 * n0 + mc, v[j][q]);
                }
            }
            fft_line<N, +1, LdsCols, TWM>(v[j], t, lds + c, TW, tw);
        }
    }
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        if (col0 + c + j * TW < inner) {
#pragma unroll
            for (int q = 0; q < P::E; ++q) {
                C *dst = XM == XM_STORE_TILES ? tile0 + (t + q * P::TPL) * TT + j * TW
                                              : base0 + (long long)(t + q * P::TPL) * inner + j * TW;
                st<NT && TC * sizeof(C) >= 128>(dst, v[j][q]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Power spectrum (src/spectrum/spectrum.cxx:63-155) binned inside the last
// forward pass of the spectrum's transform, whose output is then never
// written: each mode adds w |X|^2 (w = 1 for k_last in {0, N/2}, else 2) to
// the bin of its integer shell round(|n|).  After the transform one more LDS
// exchange gives each thread E consecutive frequencies of its line; along
// them |n| changes slowly, so the thread sums runs of equal bins in a
// register and adds each run to the block's FP64 histogram in LDS with one
// atomic (a flush whenever the bin changes, correct in any order).  Each
// block writes its histogram to partials[block][nbins], which bin_reduce
// sums in a fixed order.
// ---------------------------------------------------------------------------
struct BinArgs {
    int nbins, spb;          // bin j gathers the shells [j spb, (j+1) spb)
    double *partials;        // [gridDim.x][nbins]
};

// round(sqrt(m)) for 0 <= m < 2^24, in FP32 where m, s^2, s^2 + s and
// (s + 1)^2 are all exact: the hardware square root is within one of the
// integer square root s, which one step each way makes exact; round up
// where m > s^2 + s, i.e. m > (s + 1/2)^2
__device__ __forceinline__ int shell_of(int m)
{
    const float mf = (float)m;
    float s = floorf(__builtin_amdgcn_sqrtf(mf));
    s = (s * s > mf) ? s - 1.0f : s;
    s = ((s + 1.0f) * (s + 1.0f) <= mf) ? s + 1.0f : s;
    return (int)((mf > fmaf(s, s, s)) ? s + 1.0f : s);
}

// the block's histogram and the bin of each shell (entries up to shell
// N/2 + 1; shells beyond N/2, and beyond the last bin, map to -1)
template <int N> __device__ __forceinline__ void bin_tables(double *hist, short *sbin, const BinArgs &ba)
{
    for (int i = threadIdx.x; i < ba.nbins; i += blockDim.x) hist[i] = 0.0;
    for (int s = threadIdx.x; s <= N / 2 + 1; s += blockDim.x)
        sbin[s] = (short)((s <= N / 2 && s / ba.spb < ba.nbins) ? s / ba.spb : -1);
}
// next shell along a line: |n| changes by at most one per step, so the
// shell s of the previous position moves by at most one (FP32 exact, as in
// shell_of)
__device__ __forceinline__ float shell_step(float s, int m)
{
    const float mf = (float)m;
    s = (mf > fmaf(s, s, s)) ? s + 1.0f : s;
    return (mf <= fmaf(s, s, -s) && s > 0.0f) ? s - 1.0f : s;
}

// this thread's E consecutive frequencies p0 .. p0 + E - 1 of one line
// (m0 = the squared frequencies of the other axes, w the Hermitian weight)
template <int N, int E, typename C>
__device__ __forceinline__ void bin_run(double *hist, const short *sbin, const C (&x)[E], int p0, int m0, double w)
{
    int cur = -1;
    double acc = 0.0;
    float sh = 0.0f;
#pragma unroll
    for (int j = 0; j < E; ++j) {
        const int n = foldi(p0 + j, N);
        sh = j == 0 ? (float)shell_of(n * n + m0) : shell_step(sh, n * n + m0);
        const int bin = sbin[min((int)sh, N / 2 + 1)];
        if (bin != cur) {
            if (cur >= 0) atomicAdd(&hist[cur], acc);
            cur = bin;
            acc = 0.0;
        }
        acc = fma(w, fma((double)x[j].x, (double)x[j].x, (double)x[j].y * (double)x[j].y), acc);
    }
    if (cur >= 0) atomicAdd(&hist[cur], acc);
}
template <int N>
__device__ __forceinline__ void bin_flush(const double *hist, const BinArgs &ba)
{
    __syncthreads();
    for (int i = threadIdx.x; i < ba.nbins; i += blockDim.x) ba.partials[(long long)blockIdx.x * ba.nbins + i] = hist[i];
}

// Lines along the contiguous axis of the half spectrum, every other axis
// already transformed: 3D vt [k2][i0][k1] (D3; line = k2 N + i0), 2D vt
// [k1][i0] (line = k1).  Each wave transforms its own line (wave-local LDS
// exchanges: no block barrier inside a line), then bins it; grid-stride
// over blocks of LPB lines, with the next block's loads issued (FP32) once
// the current lines are transformed, so that they are in flight while those
// are binned.
template <int N, typename R, bool D3>
__global__ void __launch_bounds__((Geom<N, R>::CONTIG_THREADS))
contig_bin(const typename cplx<R>::T *__restrict__ data, long long nlines, BinArgs ba,
           const typename cplx<R>::T *__restrict__ twg)
{
    using C = typename cplx<R>::T;
    using P = typename Geom<N, R>::PR;
    using G_ = Geom<N, R>;
    __shared__ C lds[G_::LPB * P::PADDED];
    __shared__ C tw[N];
    __shared__ double hist[N / 2 + 1];
    __shared__ short sbin[N / 2 + 2];
    load_tw<N>(tw, twg);
    bin_tables<N>(hist, sbin, ba);
    __syncthreads();
    const int t = threadIdx.x % P::TPL, l = threadIdx.x / P::TPL;
    C *mine = lds + l * P::PADDED;
    const long long nlb = (nlines + G_::LPB - 1) / G_::LPB;
    constexpr bool PF = std::is_same<R, float>::value;
    auto load = [&](long long lb, C (&v)[P::E]) {
        const long long line = lb * G_::LPB + l;
        const bool ok = line < nlines;
        const C *src = data + line * N + t;
#pragma unroll
        for (int q = 0; q < P::E; ++q) v[q] = ok ? src[q * P::TPL] : C{0, 0};
    };
    C v[P::E];
    if (PF && blockIdx.x < nlb) load(blockIdx.x, v);
    for (long long lb = blockIdx.x; lb < nlb; lb += gridDim.x) {
        const long long line = lb * G_::LPB + l;
        if (!PF) load(lb, v);
        fft_line<N, -1, LdsContig>(v, t, mine, 1, tw);
        // positions E t .. E t + E - 1 of this line
#pragma unroll
        for (int q = 0; q < P::E; ++q) mine[LdsContig::at(t + q * P::TPL, 1)] = v[q];
        line_sync<N, LdsContig, P::E>();
        C x[P::E];
#pragma unroll
        for (int j = 0; j < P::E; ++j) x[j] = mine[LdsContig::at(P::E * t + j, 1)];
        line_sync<N, LdsContig, P::E>();
        if (PF && lb + gridDim.x < nlb) load(lb + gridDim.x, v);
        if (line < nlines) {
            // the half axis' index kh (weight 1 at 0 and N/2, else 2) and the
            // other full axis' frequency (3D)
            const int kh = D3 ? (int)(line / N) : (int)line, no = D3 ? foldi((int)(line % N), N) : 0;
            bin_run<N, P::E>(hist, sbin, x, P::E * t, kh * kh + no * no, (kh == 0 || kh == N / 2) ? 1.0 : 2.0);
        }
    }
    bin_flush<N>(hist, ba);
}

// ---------------------------------------------------------------------------
// real rows -> vt (half spectrum, k-major), two rows per complex transform
// ---------------------------------------------------------------------------
enum {
    R2C_REAL = 0,      // input: real rows (vin)
    R2C_DENSITY = 1,   // input: psi rows, rho = pref (|psi|^2 - rho_bar)
    R2C_DENSITY_TILES = 3,  // as DENSITY, psi read from a 3D lattice's column
                            // tiles (the `psi` argument; see tile_index)
};

template <int N, typename R, int MODE, bool NT>
__global__ void __launch_bounds__((Geom<N, R>::ROW_THREADS))
r2c_rows(typename cplx<R>::T *__restrict__ psi, const R *__restrict__ vin,
         typename cplx<R>::T *__restrict__ vt, long long rows, long long rpm, R pref, R rho_bar,
         const typename cplx<R>::T *__restrict__ twg, unsigned *__restrict__ tick)
{
    using C = typename cplx<R>::T;
    using P = typename Geom<N, R>::PR;
    using G = Geom<N, R>;
    constexpr int TL = G::TL, H1 = N / 2 + 1;
    __shared__ C lds[G::ROW_LDS1];
    __shared__ C tw[N];
    load_tw<N>(tw, twg);
    // the density pass of a step follows the step's drift: it advances the
    // drift counter for the next drift (no pass between them reads it)
    if (tick && blockIdx.x == 0 && threadIdx.x == 0) *tick += 1u;
    const int t = threadIdx.x % P::TPL, p = threadIdx.x / P::TPL;
    const long long r0 = (G::R2C_XCD ? xcd_block() : (long long)blockIdx.x) * TL;
    const long long ra = r0 + 2 * p, rb = ra + 1;
    const bool oka = ra < rows, okb = rb < rows;
    C *mine = lds + p * P::PADDED;
    R xa[P::E], xb[P::E];
    if (MODE == R2C_REAL) {
#pragma unroll
        for (int q = 0; q < P::E; ++q) {
            const int pos = t + q * P::TPL;
            xa[q] = oka ? vin[ra * N + pos] : (R)0;
            xb[q] = okb ? vin[rb * N + pos] : (R)0;
        }
        __syncthreads();
    } else {
        C a[P::E], b[P::E];
        if constexpr (MODE == R2C_DENSITY_TILES) {
            load_tile_row<C, N, G::TT, P::TPL, NT>(a, psi, ra, t, oka);
            load_tile_row<C, N, G::TT, P::TPL, NT>(b, psi, rb, t, okb);
        } else {
#pragma unroll
            for (int q = 0; q < P::E; ++q) {
                const int pos = t + q * P::TPL;
                a[q] = oka ? ld<NT>(psi + ra * N + pos) : C{0, 0};
                b[q] = okb ? ld<NT>(psi + rb * N + pos) : C{0, 0};
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < P::E; ++q) {
            xa[q] = pref * (a[q].x * a[q].x + a[q].y * a[q].y - rho_bar);
            xb[q] = pref * (b[q].x * b[q].x + b[q].y * b[q].y - rho_bar);
        }
    }
    // z = xa + i xb, Z = FFT(z); Xa[k] = (Z[k] + conj Z[N-k]) / 2,
    // Xb[k] = (Z[k] - conj Z[N-k]) / (2i)
    C z[P::E];
#pragma unroll
    for (int q = 0; q < P::E; ++q) z[q] = C{xa[q], xb[q]};
    fft_line<N, -1, LdsContig>(z, t, mine, 1, tw);
#pragma unroll
    for (int q = 0; q < P::E; ++q) mine[LdsContig::at(t + q * P::TPL, 1)] = z[q];
    // the pair's own line: a wave-local wait where one wavefront holds it
    line_sync<N, LdsContig, P::E>();
    C oa[P::E / 2 + 1], ob[P::E / 2 + 1];
#pragma unroll
    for (int q = 0; q <= P::E / 2; ++q) {
        const int k = t + q * P::TPL;
        if (k <= N / 2) {
            const C zk = mine[LdsContig::at(k, 1)];
            const C zm = mine[LdsContig::at((N - k) & (N - 1), 1)];
            oa[q] = C{(R)0.5 * (zk.x + zm.x), (R)0.5 * (zk.y - zm.y)};
            ob[q] = C{(R)0.5 * (zk.y + zm.y), (R)0.5 * (zm.x - zk.x)};
        }
    }
    __syncthreads();
    C *stg = lds;  // staging aliases the exchange area: stg[k*STP + l]
#pragma unroll
    for (int q = 0; q <= P::E / 2; ++q) {
        const int k = t + q * P::TPL;
        if (k <= N / 2) {
            stg[k * G::STP + 2 * p] = oa[q];
            stg[k * G::STP + 2 * p + 1] = ob[q];
        }
    }
    __syncthreads();
    // where a block's segment of a k-row is half a line, plain stores: the
    // blocks holding the two halves run on one XCD (R2C_XCD), whose L2 then
    // merges them (3D 1024^3: 2.72 against 2.87 ms streamed)
    constexpr bool NTS = NT && !G::R2C_XCD;
    if (rpm % TL == 0) {
        C *const vb = vt + vt_block_base(r0, rpm, H1);
        for (int i = threadIdx.x; i < H1 * TL; i += blockDim.x) {
            const int k = i / TL, l = i % TL;
            if (r0 + l < rows) st<NTS>(vb + (long long)k * rpm + l, stg[k * G::STP + l]);
        }
    } else {
        for (int i = threadIdx.x; i < H1 * TL; i += blockDim.x) {
            const int k = i / TL, l = i % TL;
            if (r0 + l < rows) st<NTS>(vt + vt_index(r0 + l, k, rpm, H1), stg[k * G::STP + l]);
        }
    }
}

// ---------------------------------------------------------------------------
// vt (or real V) -> real rows, fused max|V|, kick and forward transform
// ---------------------------------------------------------------------------
enum {
    C2R_SRC_VT = 1,     // rows come from vt (else from the real array V)
    C2R_STORE_V = 2,    // write the real rows to V
    C2R_VMAX = 4,       // fold max|V| into *vmax_bits
    C2R_KICK = 8,       // psi <- psi exp(-i kick_dt V)
    C2R_DRIFT = 16,     // drift along the last axis: forward transform, the
                        // axis' factor ptab[k] / N, inverse transform
    C2R_SRC_TILES = 32, // psi read from a 3D lattice's column tiles (`tiles`,
                        // see tile_index); psi rows are written
};

template <int N, typename R, bool NT>
__global__ void __launch_bounds__((Geom<N, R>::ROW_THREADS_C))
__attribute__((amdgpu_waves_per_eu(Geom<N, R>::C2R_WAVES)))
c2r_rows(const typename cplx<R>::T *__restrict__ vt, R *__restrict__ V,
         typename cplx<R>::T *__restrict__ psi, long long rows, long long rpm, R kick_dt,
         const double *__restrict__ kick_member, int flags,
         unsigned long long *__restrict__ vmax_bits, const typename cplx<R>::T *__restrict__ twg,
         const typename cplx<R>::T *__restrict__ tiles, const double2 *__restrict__ ptab,
         const float2 *__restrict__ ftab, long long ptab_stride, double inv_n,
         const typename cplx<R>::T *__restrict__ twd, const unsigned *__restrict__ tick)
{
    using C = typename cplx<R>::T;
    using P = typename Geom<N, R>::PR;
    using G = Geom<N, R>;
    constexpr int TL = G::TLC, H1 = N / 2 + 1;
    __shared__ C lds[G::ROW_LDS];
    // FP32 (TwoWord): twiddles with their low words, interleaved; the
    // drift's transforms act on psi every step and take both (TW_TWO), the
    // potential's transform the rounded twiddles alone (TW_HI)
    constexpr bool XT = TwoWord<N, R>::value;
    constexpr int TWH = XT ? TW_HI : TW_PLAIN, TWD = XT ? TW_TWO : TW_PLAIN;
    __shared__ __align__(16) C tw[XT ? 2 * N : N];
    __shared__ double wmax[16];
    // longer FP32 lines: a drifting block takes the current dithered table
    // for the drift's transforms and the potential's
    if constexpr (XT) load_tw2((float2 *)tw, (const float2 *)twg, N);
    else if constexpr (Dither<N, R>::value) load_tw<N>(tw, (flags & C2R_DRIFT) ? drift_table(twd, tick, N) : twg);
    else load_tw<N>(tw, twg);
    const int t = threadIdx.x % P::TPL, p = threadIdx.x / P::TPL;
    // where a block's rows are less than a 128 B line of vt, the blocks
    // holding the rest of the line run next on the same XCD (blocks that read
    // whole lines keep their arrival order: at 512^3 that took c2r_rows from
    // 552 to 545 us in the kernel trace)
    const long long r0 = (TL * (int)sizeof(typename cplx<R>::T) < 128 ? xcd_block() : (long long)blockIdx.x) * TL;
    const long long ra = r0 + 2 * p, rb = ra + 1;
    const bool oka = ra < rows, okb = rb < rows;
    C *mine = lds + p * P::PADDED;
    // psi's rows first: their loads overlap the potential's transform
    const bool upd = (flags & (C2R_KICK | C2R_DRIFT | C2R_SRC_TILES)) != 0;
    C a[P::E], b[P::E];
    if (upd) {
        if (flags & C2R_SRC_TILES) {
            load_tile_row<C, N, G::TT, P::TPL, NT>(a, tiles, ra, t, oka);
            load_tile_row<C, N, G::TT, P::TPL, NT>(b, tiles, rb, t, okb);
        } else {
#pragma unroll
            for (int q = 0; q < P::E; ++q) {
                const int pos = t + q * P::TPL;
                a[q] = oka ? ld<NT>(psi + ra * N + pos) : C{0, 0};
                b[q] = okb ? ld<NT>(psi + rb * N + pos) : C{0, 0};
            }
        }
    }
    R va[P::E] = {}, vb[P::E] = {};
    if (flags & C2R_SRC_VT) {
        // stage [k][l] of the block's TL rows: all loads issued before the
        // first LDS store
        C *stg = lds;
        constexpr int NS = (H1 * TL + G::ROW_THREADS_C - 1) / G::ROW_THREADS_C;
        C tmp[NS];
        if (rpm % TL == 0) {
            const C *const vb = vt + vt_block_base(r0, rpm, H1);
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                const int i = threadIdx.x + j * G::ROW_THREADS_C;
                const int k = i / TL, l = i % TL;
                tmp[j] = (i < H1 * TL && r0 + l < rows) ? vb[(long long)k * rpm + l] : C{0, 0};
            }
        } else {
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                const int i = threadIdx.x + j * G::ROW_THREADS_C;
                const int k = i / TL, l = i % TL;
                tmp[j] = (i < H1 * TL && r0 + l < rows) ? vt[vt_index(r0 + l, k, rpm, H1)] : C{0, 0};
            }
        }
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            const int i = threadIdx.x + j * G::ROW_THREADS_C;
            if (i < H1 * TL) stg[(i / TL) * G::STPC + i % TL] = tmp[j];
        }
        __syncthreads();
        C z[P::E];
#pragma unroll
        for (int q = 0; q < P::E; ++q) {
            const int j = t + q * P::TPL;
            const bool lo = j <= N / 2;
            const int k = lo ? j : N - j;
            C xa = stg[k * G::STPC + 2 * p], xb = stg[k * G::STPC + 2 * p + 1];
            if (k == 0 || k == N / 2) { xa.y = 0; xb.y = 0; }
            // Z[j] = Xa[j] + i Xb[j] (j <= N/2), conj Xa[N-j] + i conj Xb[N-j]
            z[q] = lo ? C{xa.x - xb.y, xa.y + xb.x} : C{xa.x + xb.y, xb.x - xa.y};
        }
        __syncthreads();
        fft_line<N, +1, LdsContig, TWH>(z, t, mine, 1, tw);
#pragma unroll
        for (int q = 0; q < P::E; ++q) { va[q] = z[q].x; vb[q] = z[q].y; }
    } else if (flags & (C2R_STORE_V | C2R_VMAX | C2R_KICK)) {
#pragma unroll
        for (int q = 0; q < P::E; ++q) {
            const int pos = t + q * P::TPL;
            va[q] = oka ? V[ra * N + pos] : (R)0;
            vb[q] = okb ? V[rb * N + pos] : (R)0;
        }
    }
    if (flags & C2R_STORE_V) {
#pragma unroll
        for (int q = 0; q < P::E; ++q) {
            const int pos = t + q * P::TPL;
            if (oka) st<NT>(V + ra * N + pos, va[q]);
            if (okb) st<NT>(V + rb * N + pos, vb[q]);
        }
    }
    if (flags & C2R_VMAX) {
        double m = 0.0;
#pragma unroll
        for (int q = 0; q < P::E; ++q) {
            if (oka) m = fmax(m, (double)fabs(va[q]));
            if (okb) m = fmax(m, (double)fabs(vb[q]));
        }
        for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < (int)((blockDim.x + 63) >> 6); ++w) m = fmax(m, wmax[w]);
            // non-negative doubles order like their bit patterns
            atomicMax(vmax_bits, (unsigned long long)__double_as_longlong(m));
        }
    }
    if (upd) {
        if (flags & C2R_KICK) {
            const R dta = kick_member ? (R)kick_member[oka ? ra / rpm : 0] : kick_dt;
            const R dtb = kick_member ? (R)kick_member[okb ? rb / rpm : 0] : kick_dt;
            // one test for all of this thread's angles: the polynomials cover
            // |x| <= pi/4, which the method's step sizes keep to, and the
            // short ones |x| <= 2^-6 (FP32), the usual range
            R big = 0;
#pragma unroll
            for (int q = 0; q < P::E; ++q) big = fmax(big, fmax(fabs(dta * va[q]), fabs(dtb * vb[q])));
            if (std::is_same<R, float>::value && big <= (R)0.015625) {
#pragma unroll
                for (int q = 0; q < P::E; ++q) {
                    R sa, ca, sb, cb;
                    sincosm1_short(-dta * va[q], &sa, &ca);
                    sincosm1_short(-dtb * vb[q], &sb, &cb);
                    a[q] = rotate(a[q], sa, ca);
                    b[q] = rotate(b[q], sb, cb);
                }
            } else if (big <= (R)0.78539816339744831) {
#pragma unroll
                for (int q = 0; q < P::E; ++q) {
                    R sa, ca, sb, cb;
                    sincosm1_poly(-dta * va[q], &sa, &ca);
                    sincosm1_poly(-dtb * vb[q], &sb, &cb);
                    a[q] = rotate(a[q], sa, ca);
                    b[q] = rotate(b[q], sb, cb);
                }
            } else {
#pragma unroll
                for (int q = 0; q < P::E; ++q) {
                    R sa, ca, sb, cb;
                    sincosm1_kick(-dta * va[q], &sa, &ca);
                    sincosm1_kick(-dtb * vb[q], &sb, &cb);
                    a[q] = rotate(a[q], sa, ca);
                    b[q] = rotate(b[q], sb, cb);
                }
            }
        }
        if (flags & C2R_DRIFT) {
            const long long ma_ = (oka ? ra / rpm : 0) * ptab_stride, mb_ = (okb ? rb / rpm : 0) * ptab_stride;
            C *mine_b = lds + (G::NPC + p) * P::PADDED;
            // the twiddles loaded at the start are visible after the
            // staging's barrier; without staging, after this one.  A pair's
            // exchange lines are its own, and the staging area's last reads
            // precede the potential transform's barrier
            if (!(flags & C2R_SRC_VT)) __syncthreads();
            if constexpr (G::TWO_LINES) {
                fft_line2<N, -1, LdsContig, TWD>(a, b, t, mine, mine_b, 1, tw);
            } else if constexpr (G::SHARED_LOCKSTEP) {
                fft_line2s<N, -1, TWD>(a, b, t, mine, tw);
            } else {
                fft_line<N, -1, LdsContig, TWD>(a, t, mine, 1, tw);
                fft_line<N, -1, LdsContig, TWD>(b, t, mine, 1, tw);
            }
            if constexpr (std::is_same<R, float>::value) {
                const float2 *fa = ftab + ma_, *fb = ftab + mb_;
                const float s = (float)inv_n;
#pragma unroll
                for (int q = 0; q < P::E; ++q) {
                    const int pos = t + q * P::TPL;
                    a[q] = drift_axis(fa[pos], s, a[q]);
                    b[q] = drift_axis(fb[pos], s, b[q]);
                }
            } else {
                const double2 *pa = ptab + ma_, *pb = ptab + mb_;
                const double2 one{1.0, 0.0};
#pragma unroll
                for (int q = 0; q < P::E; ++q) {
                    const int pos = t + q * P::TPL;
                    a[q] = drift_factor(pa[pos], one, inv_n, a[q]);
                    b[q] = drift_factor(pb[pos], one, inv_n, b[q]);
                }
            }
            if constexpr (G::TWO_LINES) {
                fft_line2<N, +1, LdsContig, TWD>(a, b, t, mine, mine_b, 1, tw);
            } else if constexpr (G::SHARED_LOCKSTEP) {
                fft_line2s<N, +1, TWD>(a, b, t, mine, tw);
            } else {
                fft_line<N, +1, LdsContig, TWD>(a, t, mine, 1, tw);
                fft_line<N, +1, LdsContig, TWD>(b, t, mine, 1, tw);
            }
        }
#pragma unroll
        for (int q = 0; q < P::E; ++q) {
            const int pos = t + q * P::TPL;
            if (oka) st<NT>(psi + ra * N + pos, a[q]);
            if (okb) st<NT>(psi + rb * N + pos, b[q]);
        }
    }
}

}  // namespace qx
