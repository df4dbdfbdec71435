// fft_fixed.h -- Stockham FFT with a compile-time plan, in place in one LDS
// buffer.
//
// For the y lengths that occur most (2^k and 3*2^k) the radix sequence,
// butterfly counts and strides are template constants, so each stage is
// straight-line code: every thread loads all of its butterflies of the stage
// into registers, the workgroup synchronises, and the outputs go back to the
// same buffer.  One buffer instead of two halves the LDS of the generic
// ping-pong path (fft_lds.h), so more workgroups stay resident per CU.  The
// arithmetic per butterfly (twiddle product, R-point DFT) is the generic
// path's, so both paths give bit-identical transforms.
#pragma once

#include "fft_lds.h"

namespace mmc {

template <int... Rs> struct Radices {};

template <int N> struct FixedPlan;   // only the sizes listed are specialised
template <> struct FixedPlan<16> { using type = Radices<4, 4>; };
template <> struct FixedPlan<32> { using type = Radices<4, 4, 2>; };
template <> struct FixedPlan<64> { using type = Radices<4, 4, 4>; };
template <> struct FixedPlan<96> { using type = Radices<4, 4, 2, 3>; };
template <> struct FixedPlan<128> { using type = Radices<8, 4, 4>; };
template <> struct FixedPlan<192> { using type = Radices<8, 8, 3>; };
template <> struct FixedPlan<256> { using type = Radices<8, 8, 4>; };
template <> struct FixedPlan<384> { using type = Radices<8, 8, 2, 3>; };
template <> struct FixedPlan<512> { using type = Radices<8, 8, 8>; };
template <> struct FixedPlan<768> { using type = Radices<4, 4, 4, 4, 3>; };
template <> struct FixedPlan<1024> { using type = Radices<4, 4, 4, 4, 4>; };
template <> struct FixedPlan<2048> { using type = Radices<4, 4, 4, 4, 4, 2>; };

// Lengths 128-512 take radix-8 stages (fewer LDS passes and barriers: 64
// uneven films' y pass 42.4 -> 37.7 us; at 1024 the y pass with the direct z
// convolution is slower with them, 35.4 -> 40 us, so 1024 and 2048 stay
// radix-4).  The runtime-plan path (make_plan: 4s first, then 2, 3, 5, 7)
// never transforms a length that has a compile-time plan, except the FP64
// tensor, so within a context every transform of a given length has one
// arithmetic whichever kernel runs it.
__host__ __device__ constexpr bool fixed_fft_size(int n)
{
    // y and x lengths; the z pass takes 16, 32 and 64 (fixed_z_size)
    return n == 64 || n == 96 || n == 128 || n == 192 || n == 256 || n == 384 ||
           n == 512 || n == 768 || n == 1024 || n == 2048;
}

// Twiddle exp(-2 pi i t / N), t < N, from a quarter table tw[u] =
// exp(-2 pi i u / N), u < N/4, rotated by -i per quarter (exact).
template <typename T, int N>
__device__ __forceinline__ cpx<T> twiddle_q(const cpx<T> *__restrict__ tw, int t)
{
    constexpr int Q4 = N / 4;
    const int q = t / Q4, u = t - q * Q4;
    const cpx<T> w = tw[u];
    const cpx<T> r1 = {w.y, -w.x};    // -i w
    const cpx<T> r2 = {-w.x, -w.y};   // -w
    const cpx<T> r3 = {-w.y, w.x};    // i w
    return q == 0 ? w : (q == 1 ? r1 : (q == 2 ? r2 : r3));
}

// Pruned butterflies of a zero-padded transform, the same arithmetic as
// Dft<T, R> on the inputs or outputs that remain.  DftZin: the upper half of
// the inputs is zero (first stage of a transform whose upper half of input
// rows is padding), so the first butterfly level of each half is a copy.
// DftOlo: only the lower half of the outputs is needed (last stage of a
// transform whose upper half of output rows is discarded).  Both equal the
// full butterflies' values except that a zero may differ in sign.
template <typename T, int R> struct DftZin;
template <typename T> struct DftZin<T, 4> {
    static __device__ __forceinline__ void run(cpx<T> *v)
    {
        const cpx<T> b = v[0], d = v[1];
        v[0] = cadd(b, d);
        v[2] = csub(b, d);
        v[1] = {b.x + d.y, b.y - d.x};   // b - i d
        v[3] = {b.x - d.y, b.y + d.x};   // b + i d
    }
};
template <typename T> struct DftZin<T, 8> {
    static __device__ __forceinline__ void run(cpx<T> *v)
    {
        cpx<T> e[4] = {v[0], v[2], v[0], v[2]};
        cpx<T> o[4] = {v[1], v[3], v[1], v[3]};
        DftZin<T, 4>::run(e);
        DftZin<T, 4>::run(o);
        const T c = T(0.70710678118654752440);   // sqrt(2)/2
        const cpx<T> o1 = {(o[1].x + o[1].y) * c, (o[1].y - o[1].x) * c};    // o1 (1 - i)/sqrt2
        const cpx<T> o2 = {o[2].y, -o[2].x};                                  // -i o2
        const cpx<T> o3 = {(o[3].y - o[3].x) * c, -(o[3].x + o[3].y) * c};   // o3 (-1 - i)/sqrt2
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};
template <typename T, int R> struct DftOlo;
template <typename T> struct DftOlo<T, 2> {
    static __device__ __forceinline__ void run(cpx<T> *v) { v[0] = cadd(v[0], v[1]); }
};
template <typename T> struct DftOlo<T, 4> {
    static __device__ __forceinline__ void run(cpx<T> *v)
    {
        cpx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        cpx<T> c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        v[0] = cadd(a, c);
        v[1] = {b.x + d.y, b.y - d.x};   // b - i d
    }
};
template <typename T> struct DftOlo<T, 8> {
    static __device__ __forceinline__ void run(cpx<T> *v)
    {
        cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
        cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
        Dft<T, 4>::run(e);
        Dft<T, 4>::run(o);
        const T c = T(0.70710678118654752440);   // sqrt(2)/2
        const cpx<T> o1 = {(o[1].x + o[1].y) * c, (o[1].y - o[1].x) * c};    // o1 (1 - i)/sqrt2
        const cpx<T> o2 = {o[2].y, -o[2].x};                                  // -i o2
        const cpx<T> o3 = {(o[3].y - o[3].x) * c, -(o[3].x + o[3].y) * c};   // o3 (-1 - i)/sqrt2
        v[0] = cadd(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[3] = cadd(e[3], o3);
    }
};

// QTW: tw is the quarter table (twiddle_q) instead of the full one.  Element
// i of row `row` sits at s[row*RS + i*ES]: contiguous rows (ES = 1, RS = N),
// or rows side by side (ES > 1, e.g. the y columns of a z transform, RS =
// 1), whose butterflies go to neighbouring lanes row by row so that their
// LDS accesses are adjacent.
// ZIN (first stage only): inputs r >= R/2 of every butterfly, the upper half
// of the row, are zero and not read.  OLO (last stage only): outputs
// r >= R/2, the upper half of the row, are not formed or written.
template <typename T, int N, int ROWS, int NS, int R, int NT, bool QTW = false, int ES = 1,
          int RS = N, bool ZIN = false, bool OLO = false>
__device__ __forceinline__ void fixed_stage(cpx<T> *s, const cpx<T> *__restrict__ tw)
{
    constexpr int NB = N / R;
    constexpr int TOTAL = NB * ROWS;
    constexpr int PER = (TOTAL + NT - 1) / NT;
    constexpr int TWS = N / (NS * R);
    static_assert(!ZIN || (NS == 1 && (R == 4 || R == 8)), "pruned inputs: first stage, radix 4 or 8");
    static_assert(!OLO || (NS * R == N && (R == 2 || R == 4 || R == 8)), "pruned outputs: last stage, even radix");
    constexpr int RIN = ZIN ? R / 2 : R, ROUT = OLO ? R / 2 : R;
    cpx<T> v[PER][R];
    #pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int t = threadIdx.x + b * NT;
        if ((b + 1) * NT <= TOTAL || t < TOTAL) {
            const int row = ES == 1 ? t / NB : t % ROWS, j = ES == 1 ? t % NB : t / ROWS;
            #pragma unroll
            for (int r = 0; r < RIN; ++r) v[b][r] = s[row * RS + (j + r * NB) * ES];
        }
    }
    __syncthreads();
    #pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int t = threadIdx.x + b * NT;
        if ((b + 1) * NT <= TOTAL || t < TOTAL) {
            const int row = ES == 1 ? t / NB : t % ROWS, j = ES == 1 ? t % NB : t / ROWS;
            const int k = j % NS;
            if (NS > 1) {
                #pragma unroll
                for (int r = 1; r < R; ++r)
                    v[b][r] = cmul(v[b][r], QTW ? twiddle_q<T, N>(tw, r * k * TWS) : tw[r * k * TWS]);
            }
            if constexpr (ZIN)
                DftZin<T, R>::run(v[b]);
            else if constexpr (OLO)
                DftOlo<T, R>::run(v[b]);
            else
                Dft<T, R>::run(v[b]);
            const int o = (j - k) * R + k;
            #pragma unroll
            for (int r = 0; r < ROUT; ++r) s[row * RS + (o + r * NS) * ES] = v[b][r];
        }
    }
    __syncthreads();
}

template <typename T, int N, int ROWS, int NS, int NT, bool QTW, int ES, int RS, bool ZIN, bool OLO,
          typename RL>
struct FixedStages;
template <typename T, int N, int ROWS, int NS, int NT, bool QTW, int ES, int RS, bool ZIN, bool OLO>
struct FixedStages<T, N, ROWS, NS, NT, QTW, ES, RS, ZIN, OLO, Radices<>> {
    static __device__ __forceinline__ void run(cpx<T> *, const cpx<T> *) {}
};
template <typename T, int N, int ROWS, int NS, int NT, bool QTW, int ES, int RS, bool ZIN, bool OLO,
          int R, int... Rest>
struct FixedStages<T, N, ROWS, NS, NT, QTW, ES, RS, ZIN, OLO, Radices<R, Rest...>> {
    static __device__ __forceinline__ void run(cpx<T> *s, const cpx<T> *tw)
    {
        constexpr bool last = sizeof...(Rest) == 0;
        fixed_stage<T, N, ROWS, NS, R, NT, QTW, ES, RS, ZIN && NS == 1, OLO && last>(s, tw);
        FixedStages<T, N, ROWS, NS * R, NT, QTW, ES, RS, ZIN, OLO, Radices<Rest...>>::run(s, tw);
    }
};

// Forward transform of ROWS contiguous rows of length N held in s (row
// stride N) by a workgroup of NT threads, twiddles from the full table (or
// the quarter table, QTW).  ZIN: the upper half of every row is zero and
// need not be in s; OLO: only the lower half of every row's result is
// written (the plan's last radix must be even).  Caller synchronises after
// writing s; synchronised on return.
template <typename T, int N, int ROWS, int NT = kThreads, bool QTW = false, bool ZIN = false,
          bool OLO = false>
__device__ __forceinline__ void fft_fixed(cpx<T> *s, const cpx<T> *__restrict__ tw)
{
    FixedStages<T, N, ROWS, 1, NT, QTW, 1, N, ZIN, OLO, typename FixedPlan<N>::type>::run(s, tw);
}

// Forward transform along the slow axis of a [N][ROWS] tile (element i of
// row `row` at s[i*ROWS + row]): ROWS rows side by side, e.g. the y columns
// of a z transform.  Same arithmetic as fft_fixed; synchronisation alike.
template <typename T, int N, int ROWS, int NT = kThreads, bool ZIN = false, bool OLO = false>
__device__ __forceinline__ void fft_fixed_cols(cpx<T> *s, const cpx<T> *__restrict__ tw)
{
    FixedStages<T, N, ROWS, 1, NT, false, ROWS, 1, ZIN, OLO, typename FixedPlan<N>::type>::run(s, tw);
}

// z lengths with a compile-time-planned z pass (3-D films of 8-32 layers,
// or periodic z of 16, 32, 64 layers).
__host__ __device__ constexpr bool fixed_z_size(int n) { return n == 16 || n == 32 || n == 64; }

// Plans whose last radix is even, so a transform can leave out the upper
// half of its result (fft_fixed's OLO).
__host__ __device__ constexpr bool fixed_olo_ok(int n)
{
    return n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 1024 ||
           n == 2048;
}

// Columns per y-pass workgroup of the fixed path: about 3072 complex per tile.
__host__ __device__ constexpr int fixed_cols(int n) { return n >= 1024 ? 1 : 1024 / n; }

}  // namespace mmc
