// qx_fft_core.hpp -- register/LDS Stockham FFT building blocks for gfx950.
//
// A line of N complex values is spread over TPL = N/E threads; thread t holds
// the E values at positions t + q*TPL (q = 0..E-1).  Every Stockham stage
// reads its inputs in that distribution and, for the last stage, also leaves
// its outputs in it, so global loads and stores of a contiguous line are
// coalesced and a forward transform can flow straight into a pointwise
// multiply and an inverse transform without touching LDS in between.
//
// Stages (radix R, Ns = product of earlier radices) follow the self-sorting
// Stockham formulation: butterfly j takes x[j + r*N/R], twiddles by
// W_{Ns R}^{r (j mod Ns)}, and writes y[(j/Ns) Ns R + (j mod Ns) + r Ns].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace qx {

template <typename R> struct cplx;
template <> struct cplx<float> { using T = float2; };
template <> struct cplx<double> { using T = double2; };

template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * (wh + wl): the twiddle in two words, wh the rounded twiddle and wl =
// w - wh rounded, so the product carries the twiddle to ~2^-48 instead of
// the ~2^-25 of wh alone.  A rounded twiddle is the same operator error in
// every transform, and the drift applies its transforms to psi step after
// step: with one word the error of psi grows linearly with the step count;
// with two words only the (random) rounding of the arithmetic is left, which
// grows like its root.  a wl enters the fused multiply-adds that form a wh:
// added to a product rounded first, it would mostly round away (it is below
// half an ulp of it).
__device__ __forceinline__ float2 cmul2(float2 a, float2 wh, float2 wl)
{
    const float cr = fmaf(a.x, wl.x, -a.y * wl.y), ci = fmaf(a.x, wl.y, a.y * wl.x);
    return float2{fmaf(a.x, wh.x, fmaf(-a.y, wh.y, cr)), fmaf(a.x, wh.y, fmaf(a.y, wh.x, ci))};
}
__device__ __forceinline__ double2 cmul2(double2 a, double2 wh, double2) { return cmul(a, wh); }

// Twiddle tables in LDS, entry m in stage-major order (twiddle_offset).
// TW_PLAIN: tw[m], one word per twiddle.  FP32 passes that transform psi
// keep them interleaved, tw[2m] the rounded twiddle and tw[2m + 1] its low
// word (one 16-byte read per twiddle): TW_HI reads only the rounded
// twiddles of such a table, TW_TWO both words (two-word products, cmul2).
enum { TW_PLAIN = 0, TW_HI = 1, TW_TWO = 2 };
template <int TWM, int DIR, typename C>
__device__ __forceinline__ C twiddle_mul(C u, const C *tw, int m)
{
    if constexpr (TWM == TW_TWO && std::is_same<C, float2>::value) {
        const float4 w4 = reinterpret_cast<const float4 *>(tw)[m];
        const float2 wh{w4.x, DIR > 0 ? -w4.y : w4.y}, wl{w4.z, DIR > 0 ? -w4.w : w4.w};
        return cmul2(u, wh, wl);
    } else {
        C w = tw[TWM == TW_PLAIN ? m : 2 * m];
        if (DIR > 0) w.y = -w.y;
        return cmul(u, w);
    }
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }

// multiply by -i (DIR < 0) or +i (DIR > 0)
template <int DIR, typename C> __device__ __forceinline__ C mul_i(C a)
{
    return DIR < 0 ? C{a.y, -a.x} : C{-a.y, a.x};
}

// ---------------------------------------------------------------------------
// small DFTs in registers; DIR < 0 is e^{-2 pi i jk/R} (forward)
// ---------------------------------------------------------------------------
template <int R, int DIR, typename C> struct Dft;

template <int DIR, typename C> struct Dft<1, DIR, C> {
    static __device__ __forceinline__ void run(C *) {}
};
template <int DIR, typename C> struct Dft<2, DIR, C> {
    static __device__ __forceinline__ void run(C *a)
    {
        C t = a[1];
        a[1] = csub(a[0], t);
        a[0] = cadd(a[0], t);
    }
};
template <int DIR, typename C> struct Dft<4, DIR, C> {
    static __device__ __forceinline__ void run(C *a)
    {
        C t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
        C t2 = cadd(a[1], a[3]), t3 = mul_i<DIR>(csub(a[1], a[3]));
        a[0] = cadd(t0, t2);
        a[2] = csub(t0, t2);
        a[1] = cadd(t1, t3);
        a[3] = csub(t1, t3);
    }
};
// multiplication by 1/sqrt(2).  In FP32 the constant is split in two
// words: the rounded constant alone has 2h^2 = 1 - 4.1e-8, which would
// shrink every W8 rotation and bias the norm (mass) of psi step after step.
__device__ __forceinline__ float mul_rsqrt2(float a)
{
    return fmaf(0.70710676908493042f, a, 1.2101617e-08f * a);
}
__device__ __forceinline__ double mul_rsqrt2(double a) { return 0.70710678118654752440 * a; }

template <int DIR, typename C> struct Dft<8, DIR, C> {
    static __device__ __forceinline__ void run(C *a)
    {
        C e[4] = {a[0], a[2], a[4], a[6]};
        C o[4] = {a[1], a[3], a[5], a[7]};
        Dft<4, DIR, C>::run(e);
        Dft<4, DIR, C>::run(o);
        // twiddle the odd half by W8^k
        C o1, o3;
        if (DIR < 0) {
            o1 = C{mul_rsqrt2(o[1].x + o[1].y), mul_rsqrt2(o[1].y - o[1].x)};
            o3 = C{mul_rsqrt2(o[3].y - o[3].x), -mul_rsqrt2(o[3].x + o[3].y)};
        } else {
            o1 = C{mul_rsqrt2(o[1].x - o[1].y), mul_rsqrt2(o[1].x + o[1].y)};
            o3 = C{-mul_rsqrt2(o[3].x + o[3].y), mul_rsqrt2(o[3].x - o[3].y)};
        }
        C o2 = mul_i<DIR>(o[2]);
        a[0] = cadd(e[0], o[0]);
        a[4] = csub(e[0], o[0]);
        a[1] = cadd(e[1], o1);
        a[5] = csub(e[1], o1);
        a[2] = cadd(e[2], o2);
        a[6] = csub(e[2], o2);
        a[3] = cadd(e[3], o3);
        a[7] = csub(e[3], o3);
    }
};

// ---------------------------------------------------------------------------
// compile-time plan: radix-8 stages, the last one of radix 2, 4 or 8.  The
// small radix goes last: a stage's butterfly j writes positions
// (j/Ns) Ns R + (j mod Ns) + r Ns, which for a first stage of radix 2 or 4
// (Ns = 1) are 2 or 4 slots apart across lanes and hit few LDS banks; as the
// last stage it writes nothing to LDS (its outputs stay in registers)
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// the stage structure of an N-point transform, shared by host and device
__host__ __device__ constexpr int ipow8(int k) { return k == 0 ? 1 : 8 * ipow8(k - 1); }
__host__ __device__ constexpr int plan_stages(int N) { return (ilog2(N) + 2) / 3; }
__host__ __device__ constexpr int plan_r0(int N) { return (ilog2(N) % 3 == 0) ? 8 : ((ilog2(N) % 3 == 1) ? 2 : 4); }
__host__ __device__ constexpr int plan_radix(int N, int s) { return N < 8 ? N : (s == plan_stages(N) - 1 ? plan_r0(N) : 8); }
__host__ __device__ constexpr int plan_ns(int N, int s) { return (s == 0 || N < 8) ? 1 : ipow8(s); }

// Stage-major twiddle table: stage s (sub-transform length Ns > 1, radix R,
// table stride TWS = N/(Ns R)) keeps its twiddles w^{r k TWS}, r = 1..R-1,
// k = 0..Ns-1, at twiddle_offset(N, s) + (r - 1) Ns + k, with w = e^{-2 pi
// i/N}.  The lanes of a wavefront then read consecutive entries (k follows
// the thread index), free of LDS bank conflicts, where the natural order
// w^m put the reads of a stage r k TWS entries apart.  The table holds
// N - 8 entries (N >= 8).
__host__ __device__ constexpr int twiddle_offset(int N, int s)
{
    int off = 0;
    for (int i = 0; i < s; ++i)
        if (plan_ns(N, i) > 1) off += (plan_radix(N, i) - 1) * plan_ns(N, i);
    return off;
}

// EV values per thread: 8 by default (16 for the long contiguous lines of
// the row kernels, so that a line of 1024 points is one wavefront)
template <int N, int EV = (N < 8 ? N : 8)> struct Plan {
    static constexpr int E = EV;                 // values per thread
    static constexpr int TPL = N / E;            // threads per line
    static constexpr int LG = ilog2(N);
    static constexpr int NST = plan_stages(N);   // number of stages
    static constexpr int R0 = plan_r0(N);
    static constexpr int radix(int s) { return plan_radix(N, s); }
    static constexpr int ns(int s) { return plan_ns(N, s); }
    // padded LDS slots for one contiguous line (one pad slot per 8 values)
    static constexpr int PADDED = N + N / 8 + 1;
};

// LDS addressing of line positions
struct LdsContig {  // [line][pos] with one pad slot per 8 values
    __device__ __forceinline__ static int at(int pos, int) { return pos + (pos >> 3); }
};
struct LdsCols {    // [pos][column], columns interleaved
    __device__ __forceinline__ static int at(int pos, int tc) { return pos * tc; }
};

// Synchronisation of a line's LDS exchange.  A contiguous line is held by
// TPL consecutive threads; with TPL <= 64 those are lanes of one wavefront,
// whose LDS operations complete in issue order, so waiting for this wave's
// own LDS operations (and keeping the compiler from moving memory accesses
// across the wait) is enough.  Column-interleaved lines span wavefronts and
// take a block barrier.
template <int N, typename L, int EV> __device__ __forceinline__ void line_sync()
{
    if constexpr (std::is_same<L, LdsContig>::value && Plan<N, EV>::TPL <= 64)
        __asm__ __volatile__("s_waitcnt lgkmcnt(0)" ::: "memory");
    else
        __syncthreads();
}

// One Stockham stage.  v holds x[t + q TPL]; on exit (not last) v holds the
// next stage's inputs, read back from LDS; (last) v holds y[t + q TPL].
// TWM: the twiddle table's layout and use (TW_PLAIN, TW_HI, TW_TWO).
template <int N, int S, int DIR, typename L, int TWM, typename C, int EV>
__device__ __forceinline__ void fft_stage(C (&v)[EV], int t, C *line, int tc, const C *tw)
{
    using P = Plan<N, EV>;
    constexpr int E = P::E, TPL = P::TPL;
    constexpr int R = P::radix(S), Ns = P::ns(S), NB = E / R;
    constexpr bool last = (S == P::NST - 1);
    constexpr int TWO = twiddle_offset(N, S);   // this stage's table
    C out[E];
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        const int j = t + m * TPL;
        const int k = j & (Ns - 1);
        C u[R];
#pragma unroll
        for (int r = 0; r < R; ++r) u[r] = v[m + r * NB];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) u[r] = twiddle_mul<TWM, DIR>(u[r], tw, TWO + (r - 1) * Ns + k);
        }
        Dft<R, DIR, C>::run(u);
        if (last) {
#pragma unroll
            for (int r = 0; r < R; ++r) out[m + r * NB] = u[r];
        } else {
            const int base = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) line[L::at(base + r * Ns, tc)] = u[r];
        }
    }
    if (last) {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = out[q];
    } else {
        line_sync<N, L, EV>();
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = line[L::at(t + q * TPL, tc)];
        line_sync<N, L, EV>();
    }
}

template <int N, int S, int DIR, typename L, int TWM, typename C, int EV>
__device__ __forceinline__ void stage_loop(C (&v)[EV], int t, C *line, int tc, const C *tw)
{
    if constexpr (S < Plan<N, EV>::NST) {
        fft_stage<N, S, DIR, L, TWM>(v, t, line, tc, tw);
        stage_loop<N, S + 1, DIR, L, TWM>(v, t, line, tc, tw);
    }
}

// Full N-point transform of the line held in v.  All threads of the block
// must call it (column lines and lines of more than 64 threads synchronise
// with block barriers).  `line` is this line's LDS
// area, `tc` the column interleave for LdsCols, `tw` the LDS twiddle table
// in stage-major order (TW_PLAIN), or that table interleaved with the
// twiddles' low words (TW_HI, TW_TWO).
template <int N, int DIR, typename L, int TWM = TW_PLAIN, typename C, int EV>
__device__ __forceinline__ void fft_line(C (&v)[EV], int t, C *line, int tc, const C *tw)
{
    stage_loop<N, 0, DIR, L, TWM>(v, t, line, tc, tw);
}

// Two lines transformed in lockstep (same direction): each stage's
// butterflies of both lines, one synchronisation per exchange for both, two
// independent dependency chains for the scheduler.  line_a / line_b are the
// two lines' LDS areas.
template <int N, int S, int DIR, typename L, int TWM, typename C, int EV>
__device__ __forceinline__ void fft_stage2(C (&va)[EV], C (&vb)[EV], int t, C *line_a,
                                           C *line_b, int tc, const C *tw)
{
    using P = Plan<N, EV>;
    constexpr int E = P::E, TPL = P::TPL;
    constexpr int R = P::radix(S), Ns = P::ns(S), NB = E / R;
    constexpr bool last = (S == P::NST - 1);
    constexpr int TWO = twiddle_offset(N, S);
    C oa[E], ob[E];
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        const int j = t + m * TPL;
        const int k = j & (Ns - 1);
        C ua[R], ub[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            ua[r] = va[m + r * NB];
            ub[r] = vb[m + r * NB];
        }
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                ua[r] = twiddle_mul<TWM, DIR>(ua[r], tw, TWO + (r - 1) * Ns + k);
                ub[r] = twiddle_mul<TWM, DIR>(ub[r], tw, TWO + (r - 1) * Ns + k);
            }
        }
        Dft<R, DIR, C>::run(ua);
        Dft<R, DIR, C>::run(ub);
        if (last) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                oa[m + r * NB] = ua[r];
                ob[m + r * NB] = ub[r];
            }
        } else {
            const int base = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                line_a[L::at(base + r * Ns, tc)] = ua[r];
                line_b[L::at(base + r * Ns, tc)] = ub[r];
            }
        }
    }
    if (last) {
#pragma unroll
        for (int q = 0; q < E; ++q) {
            va[q] = oa[q];
            vb[q] = ob[q];
        }
    } else {
        line_sync<N, L, EV>();
#pragma unroll
        for (int q = 0; q < E; ++q) {
            va[q] = line_a[L::at(t + q * TPL, tc)];
            vb[q] = line_b[L::at(t + q * TPL, tc)];
        }
        line_sync<N, L, EV>();
    }
}

template <int N, int S, int DIR, typename L, int TWM, typename C, int EV>
__device__ __forceinline__ void stage_loop2(C (&va)[EV], C (&vb)[EV], int t, C *line_a,
                                            C *line_b, int tc, const C *tw)
{
    if constexpr (S < Plan<N, EV>::NST) {
        fft_stage2<N, S, DIR, L, TWM>(va, vb, t, line_a, line_b, tc, tw);
        stage_loop2<N, S + 1, DIR, L, TWM>(va, vb, t, line_a, line_b, tc, tw);
    }
}

template <int N, int DIR, typename L, int TWM = TW_PLAIN, typename C, int EV>
__device__ __forceinline__ void fft_line2(C (&va)[EV], C (&vb)[EV], int t, C *line_a,
                                          C *line_b, int tc, const C *tw)
{
    stage_loop2<N, 0, DIR, L, TWM>(va, vb, t, line_a, line_b, tc, tw);
}

// Two lines in lockstep through ONE exchange line: each stage's butterflies
// of both lines (each twiddle read from LDS serves both), then line a's
// exchange and line b's one after the other through the same LDS area.  The
// LDS footprint is one line's; the twiddle reads and their waits halve.
template <int N, int S, int DIR, int TWM, typename C, int EV>
__device__ __forceinline__ void fft_stage2s(C (&va)[EV], C (&vb)[EV], int t, C *line, const C *tw)
{
    using P = Plan<N, EV>;
    using L = LdsContig;
    constexpr int E = P::E, TPL = P::TPL;
    constexpr int R = P::radix(S), Ns = P::ns(S), NB = E / R;
    constexpr bool last = (S == P::NST - 1);
    constexpr int TWO = twiddle_offset(N, S);
    C oa[E], ob[E];
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        const int j = t + m * TPL;
        const int k = j & (Ns - 1);
        C ua[R], ub[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            ua[r] = va[m + r * NB];
            ub[r] = vb[m + r * NB];
        }
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                ua[r] = twiddle_mul<TWM, DIR>(ua[r], tw, TWO + (r - 1) * Ns + k);
                ub[r] = twiddle_mul<TWM, DIR>(ub[r], tw, TWO + (r - 1) * Ns + k);
            }
        }
        Dft<R, DIR, C>::run(ua);
        Dft<R, DIR, C>::run(ub);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            oa[m + r * NB] = ua[r];
            ob[m + r * NB] = ub[r];
        }
    }
    if (last) {
#pragma unroll
        for (int q = 0; q < E; ++q) {
            va[q] = oa[q];
            vb[q] = ob[q];
        }
    } else {
        // outputs of butterfly m, digit r, go to position base(m) + r Ns
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            const int j = t + m * TPL;
            const int base = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
            for (int r = 0; r < R; ++r) line[L::at(base + r * Ns, 1)] = oa[m + r * NB];
        }
        line_sync<N, L, EV>();
#pragma unroll
        for (int q = 0; q < E; ++q) va[q] = line[L::at(t + q * TPL, 1)];
        line_sync<N, L, EV>();
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            const int j = t + m * TPL;
            const int base = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
            for (int r = 0; r < R; ++r) line[L::at(base + r * Ns, 1)] = ob[m + r * NB];
        }
        line_sync<N, L, EV>();
#pragma unroll
        for (int q = 0; q < E; ++q) vb[q] = line[L::at(t + q * TPL, 1)];
        line_sync<N, L, EV>();
    }
}

template <int N, int S, int DIR, int TWM, typename C, int EV>
__device__ __forceinline__ void stage_loop2s(C (&va)[EV], C (&vb)[EV], int t, C *line, const C *tw)
{
    if constexpr (S < Plan<N, EV>::NST) {
        fft_stage2s<N, S, DIR, TWM>(va, vb, t, line, tw);
        stage_loop2s<N, S + 1, DIR, TWM>(va, vb, t, line, tw);
    }
}

// contiguous lines only (one LDS line shared by the two)
template <int N, int DIR, int TWM = TW_PLAIN, typename C, int EV>
__device__ __forceinline__ void fft_line2s(C (&va)[EV], C (&vb)[EV], int t, C *line, const C *tw)
{
    stage_loop2s<N, 0, DIR, TWM>(va, vb, t, line, tw);
}

}  // namespace qx
