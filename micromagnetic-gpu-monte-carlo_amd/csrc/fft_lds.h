// fft_lds.h -- LDS-resident Stockham FFT for one workgroup.
//
// A workgroup transforms `nrows` rows of length n that sit in LDS, element
// (row, i) at s[row*rs + i*es].  Each stage of radix R reads its R inputs
// (stride n/R) from one LDS buffer, applies the twiddles
// exp(-2 pi i r k / (Ns R)), runs the R-point DFT in registers and writes
// the outputs autosorted (Stockham) into the other buffer, so the result ends
// in natural order with no bit reversal and one barrier per stage.  Keeping
// only one butterfly in registers at a time (rather than all of a thread's
// butterflies across a barrier) holds the kernels near 50 VGPRs (FP32) with
// any mix of radices.  Twiddles come from a table tw[t] = exp(-2 pi i t / n).
// Forward transforms only; an inverse is conj -> forward -> conj, which the
// callers fold into their loads and stores.
#pragma once

#include "mmc_internal.h"

namespace mmc {

// a*b + c rounded once.  The build forms no fused multiply-add by itself
// (-ffp-contract=off), so an expression rounds the same in every kernel it
// is inlined into; these explicit ones keep that property and save an
// instruction each.
__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return {fmad(a.x, b.x, -(a.y * b.y)), fmad(a.x, b.y, a.y * b.x)};
}
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }

// acc + m0 k0 + m1 k1 + m2 k2 (complex): each part a chain of fused
// multiply-adds (a row of the symmetric 3x3 tensor product in k-space).
template <typename T>
__device__ __forceinline__ cpx<T> cdot3_acc(cpx<T> acc, cpx<T> m0, cpx<T> k0, cpx<T> m1, cpx<T> k1,
                                            cpx<T> m2, cpx<T> k2)
{
    T re = fmad(m0.x, k0.x, acc.x), im = fmad(m0.x, k0.y, acc.y);
    re = fmad(-m0.y, k0.y, re); im = fmad(m0.y, k0.x, im);
    re = fmad(m1.x, k1.x, re);  im = fmad(m1.x, k1.y, im);
    re = fmad(-m1.y, k1.y, re); im = fmad(m1.y, k1.x, im);
    re = fmad(m2.x, k2.x, re);  im = fmad(m2.x, k2.y, im);
    re = fmad(-m2.y, k2.y, re); im = fmad(m2.y, k2.x, im);
    return {re, im};
}
// m0 k0 + m1 k1 + m2 k2
template <typename T>
__device__ __forceinline__ cpx<T> cdot3(cpx<T> m0, cpx<T> k0, cpx<T> m1, cpx<T> k1, cpx<T> m2,
                                        cpx<T> k2)
{
    T re = m0.x * k0.x, im = m0.x * k0.y;
    re = fmad(-m0.y, k0.y, re); im = fmad(m0.y, k0.x, im);
    re = fmad(m1.x, k1.x, re);  im = fmad(m1.x, k1.y, im);
    re = fmad(-m1.y, k1.y, re); im = fmad(m1.y, k1.x, im);
    re = fmad(m2.x, k2.x, re);  im = fmad(m2.x, k2.y, im);
    re = fmad(-m2.y, k2.y, re); im = fmad(m2.y, k2.x, im);
    return {re, im};
}
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cconj(cpx<T> a) { return {a.x, -a.y}; }

// ---- R-point forward DFTs (sign -1) in registers ---------------------------
template <typename T, int R> struct Dft;

template <typename T> struct Dft<T, 2> {
    static __device__ __forceinline__ void run(cpx<T> *v)
    {
        cpx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename T> struct Dft<T, 3> {
    static __device__ __forceinline__ void run(cpx<T> *v)
    {
        const T c = T(-0.5);
        const T s = T(-0.86602540378443864676);   // -sin(2 pi / 3)
        cpx<T> t1 = cadd(v[1], v[2]);
        cpx<T> t2 = csub(v[1], v[2]);
        cpx<T> m = {fmad(c, t1.x, v[0].x), fmad(c, t1.y, v[0].y)};
        v[0] = cadd(v[0], t1);
        // +- i s t2
        v[1] = {fmad(-s, t2.y, m.x), fmad(s, t2.x, m.y)};
        v[2] = {fmad(s, t2.y, m.x), fmad(-s, t2.x, m.y)};
    }
};

template <typename T> struct Dft<T, 4> {
    static __device__ __forceinline__ void run(cpx<T> *v)
    {
        cpx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        cpx<T> c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = {b.x + d.y, b.y - d.x};   // b - i d
        v[3] = {b.x - d.y, b.y + d.x};   // b + i d
    }
};

// 8 points as two 4-point DFTs of the even and odd inputs joined by the
// twiddles exp(-2 pi i k/8) (exact rotations by -i, and (1 -+ i)/sqrt2).
template <typename T> struct Dft<T, 8> {
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
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

template <typename T> struct Dft<T, 5> {
    static __device__ __forceinline__ void run(cpx<T> *v)
    {
        const T c1 = T(0.30901699437494742410);    // cos(2 pi/5)
        const T c2 = T(-0.80901699437494742410);   // cos(4 pi/5)
        const T s1 = T(-0.95105651629515357212);   // -sin(2 pi/5)
        const T s2 = T(-0.58778525229247312917);   // -sin(4 pi/5)
        cpx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        cpx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        cpx<T> x0 = v[0];
        cpx<T> r1 = {fmad(c2, a2.x, fmad(c1, a1.x, x0.x)), fmad(c2, a2.y, fmad(c1, a1.y, x0.y))};
        cpx<T> r2 = {fmad(c1, a2.x, fmad(c2, a1.x, x0.x)), fmad(c1, a2.y, fmad(c2, a1.y, x0.y))};
        cpx<T> i1 = {fmad(s1, b1.x, s2 * b2.x), fmad(s1, b1.y, s2 * b2.y)};
        cpx<T> i2 = {fmad(s2, b1.x, -(s1 * b2.x)), fmad(s2, b1.y, -(s1 * b2.y))};
        v[0] = {x0.x + a1.x + a2.x, x0.y + a1.y + a2.y};
        v[1] = {r1.x - i1.y, r1.y + i1.x};   // r1 + i*i1
        v[4] = {r1.x + i1.y, r1.y - i1.x};   // r1 - i*i1
        v[2] = {r2.x - i2.y, r2.y + i2.x};
        v[3] = {r2.x + i2.y, r2.y - i2.x};
    }
};

template <typename T> struct Dft<T, 7> {
    static __device__ __forceinline__ void run(cpx<T> *v)
    {
        const T c[4] = {T(1.0), T(0.62348980185873353053), T(-0.22252093395631440429),
                        T(-0.90096886790241912624)};
        const T s[4] = {T(0.0), T(-0.78183148246802980871), T(-0.97492791218182360702),
                        T(-0.43388373911755812048)};
        cpx<T> a[4], b[4];
        #pragma unroll
        for (int r = 1; r < 4; ++r) { a[r] = cadd(v[r], v[7 - r]); b[r] = csub(v[r], v[7 - r]); }
        cpx<T> x0 = v[0];
        cpx<T> out[7];
        out[0] = {x0.x + a[1].x + a[2].x + a[3].x, x0.y + a[1].y + a[2].y + a[3].y};
        #pragma unroll
        for (int k = 1; k < 4; ++k) {
            cpx<T> re = x0, im = {T(0), T(0)};
            #pragma unroll
            for (int r = 1; r < 4; ++r) {
                const int e = (r * k) % 7;
                const int q = e < 4 ? e : 7 - e;
                const T cs = c[q];
                const T sn = e < 4 ? s[q] : -s[q];
                re.x = fmad(cs, a[r].x, re.x); re.y = fmad(cs, a[r].y, re.y);
                im.x = fmad(sn, b[r].x, im.x); im.y = fmad(sn, b[r].y, im.y);
            }
            out[k] = {re.x - im.y, re.y + im.x};
            out[7 - k] = {re.x + im.y, re.y - im.x};
        }
        #pragma unroll
        for (int r = 0; r < 7; ++r) v[r] = out[r];
    }
};

// ---- one Stockham stage: in -> out ---------------------------------------------
template <typename T, int R>
__device__ __forceinline__ void fft_stage(const cpx<T> *__restrict__ in,
                                          cpx<T> *__restrict__ out, int n, int nrows,
                                          int es, int rs, int Ns,
                                          const cpx<T> *__restrict__ tw)
{
    const int nb = n / R;
    const int total = nb * nrows;
    const int tws = n / (Ns * R);
    // power-of-two butterfly counts (every radix-4/2 plan): shifts and masks
    // instead of integer division
    const bool p2 = ((nb & (nb - 1)) == 0) && ((Ns & (Ns - 1)) == 0);
    const int lnb = __builtin_ctz((unsigned)nb);
    // Strided elements (es > 1, rows side by side: the z pass's y columns):
    // neighbouring lanes take neighbouring rows, so their LDS accesses are
    // adjacent; with butterflies along a row on neighbouring lanes they would
    // be es apart, all in one bank.  Same butterflies, other lanes.
    const bool across = es != 1;
    const bool p2r = (nrows & (nrows - 1)) == 0;
    const int lnr = __builtin_ctz((unsigned)nrows);
    for (int t = threadIdx.x; t < total; t += kThreads) {
        int row, j, k;
        if (across) {
            if (p2r) {
                row = t & (nrows - 1);
                j = t >> lnr;
            } else {
                j = t / nrows;
                row = t - j * nrows;
            }
            k = p2 ? (j & (Ns - 1)) : j % Ns;
        } else if (p2) {
            row = t >> lnb;
            j = t & (nb - 1);
            k = j & (Ns - 1);
        } else {
            row = t / nb;
            j = t - row * nb;
            k = j % Ns;
        }
        const cpx<T> *ib = in + row * rs;
        cpx<T> v[R];
        #pragma unroll
        for (int r = 0; r < R; ++r) v[r] = ib[(j + r * nb) * es];
        if (Ns > 1) {
            #pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[r * k * tws]);
        }
        Dft<T, R>::run(v);
        cpx<T> *ob = out + row * rs;
        const int o = (j - k) * R + k;
        #pragma unroll
        for (int r = 0; r < R; ++r) ob[(o + r * Ns) * es] = v[r];
    }
    __syncthreads();
}

// Transform nrows rows held in s0 with plan pl, ping-ponging with s1 (same
// size and layout).  Returns the buffer holding the result.  The caller must
// have synchronised after writing s0; the result is synchronised on return.
template <typename T>
__device__ __forceinline__ cpx<T> *fft_rows(cpx<T> *s0, cpx<T> *s1, int nrows, int es, int rs,
                            const FftPlan &pl, const cpx<T> *__restrict__ tw)
{
    const int n = pl.n;
    const int nst = pl.nst;
    int Ns = 1;
    for (int st = 0; st < nst; ++st) {
        const int R = pl.radix[st];
        switch (R) {
        case 4: fft_stage<T, 4>(s0, s1, n, nrows, es, rs, Ns, tw); break;
        case 2: fft_stage<T, 2>(s0, s1, n, nrows, es, rs, Ns, tw); break;
        case 3: fft_stage<T, 3>(s0, s1, n, nrows, es, rs, Ns, tw); break;
        case 5: fft_stage<T, 5>(s0, s1, n, nrows, es, rs, Ns, tw); break;
        default: fft_stage<T, 7>(s0, s1, n, nrows, es, rs, Ns, tw); break;
        }
        cpx<T> *t = s0;
        s0 = s1;
        s1 = t;
        Ns *= R;
    }
    return s0;
}

}  // namespace mmc
