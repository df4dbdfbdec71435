// k_xpass.hip -- x transforms: real rows <-> half spectrum along x, two
// real rows packed in one complex row, stored transposed so y is contiguous
// (generic runtime-plan kernels and compile-time-planned ones per length).
#include "mmc_device.h"

namespace mmc {

// mode 0: magnetisation (m -> A), extents (nx, ny, nz), 3 components
// mode 1: tensor (K -> KA, FP64), extents (Px, Py, Pz), 6 components
template <typename T>
__global__ __launch_bounds__(kThreads) void k_xfwd(const Inst *__restrict__ I,
                                                   const Work *__restrict__ W, int mode,
                                                   const T *__restrict__ realb,
                                                   cpx<T> *__restrict__ specb,
                                                   const cpx<T> *__restrict__ twb)
{
    cpx<T> *s = reinterpret_cast<cpx<T> *>(g_smem);
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const bool p16 = pair16<T>(in);
    const int c = w.a, z = w.b, y0 = w.c;
    const int P = in.Px, H = in.Hx;
    int Lx, Ly, Lz, ty;
    const T *src;
    cpx<T> *dst;
    if (mode == 0) {
        Lx = in.nx; Ly = in.ny; Lz = in.nz; ty = in.tyx;
        src = realb + in.m_off + (long)c * in.N;
        dst = specb + in.a_off + (long)c * H * Lz * Ly;
    } else {
        Lx = in.Px; Ly = in.Py; Lz = in.Pz; ty = in.tyx64;
        src = realb + in.k_off + (long)c * in.NP;
        dst = specb + in.ka_off + (long)c * H * Lz * Ly;
    }
    src += (long)z * Lx * Ly;
    cpx<T> *s1 = s + (ty >> 1) * P;
    const cpx<T> *tw = lds_twiddles(s1 + (ty >> 1) * P, twb + in.twx, P);
    const int nyt = min(ty, Ly - y0);
    const int npair = (nyt + 1) >> 1;
    staged_copy_dyn<cpx<T>>(
        npair * P,
        [&](int e) {
            const int q = e / P, x = e - q * P;
            T re = T(0), im = T(0);
            if (x < Lx) {
                const int ya = y0 + 2 * q;
                re = src[(long)ya * Lx + x];
                if (2 * q + 1 < nyt) im = src[(long)(ya + 1) * Lx + x];
            }
            return cpx<T>{re, im};
        },
        [&](int e, cpx<T> v) { s[e] = v; });
    __syncthreads();
    s = fft_rows<T>(s, s1, npair, 1, P, in.plx, tw);
    // split the packed spectrum: Xa = (Z_k + conj Z_-k)/2, Xb = -i(Z_k - conj Z_-k)/2
    for (int e = threadIdx.x; e < H * npair; e += kThreads) {
        const int k = e / npair, q = e - k * npair;
        const cpx<T> A = s[q * P + k];
        const cpx<T> B = s[q * P + (k ? P - k : 0)];
        cpx<T> *o = dst + ((long)k * Lz + z) * Ly + y0 + 2 * q;
        const cpx<T> oa = {(A.x + B.x) * T(0.5), (A.y - B.y) * T(0.5)};
        const cpx<T> ob = {(A.y + B.y) * T(0.5), (B.x - A.x) * T(0.5)};
        if (p16) {
            store_pair16(o, oa, ob);
        } else {
            o[0] = oa;
            if (2 * q + 1 < nyt) o[1] = ob;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_xinv(const Inst *__restrict__ I,
                                                   const Work *__restrict__ W,
                                                   const cpx<T> *__restrict__ specb,
                                                   T *__restrict__ realb,
                                                   const cpx<T> *__restrict__ twb)
{
    cpx<T> *s = reinterpret_cast<cpx<T> *>(g_smem);
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const bool p16 = pair16<T>(in);
    const int c = w.a, z = w.b, y0 = w.c;
    const int P = in.Px, H = in.Hx;
    const int Lx = in.nx, Ly = in.ny, Lz = in.nz;
    cpx<T> *s1 = s + (in.tyx >> 1) * P;
    const cpx<T> *tw = lds_twiddles(s1 + (in.tyx >> 1) * P, twb + in.twx, P);
    const cpx<T> *src = specb + in.a_off + (long)c * H * Lz * Ly;
    T *dst = realb + in.h_off + (long)c * in.N + (long)z * Lx * Ly;
    const int nyt = min(in.tyx, Ly - y0);
    const int npair = (nyt + 1) >> 1;
    // Z = Xa + i Xb on k < H, Hermitian completion above; stored conjugated
    // so the forward transform computes the inverse.
    staged_copy_dyn<Pair2<cpx<T>>>(
        H * npair,
        [&](int e) {
            const int k = e / npair, q = e - k * npair;
            const cpx<T> *p = src + ((long)k * Lz + z) * Ly + y0 + 2 * q;
            if (p16) return load_pair16(p);
            Pair2<cpx<T>> v;
            // branch-free: p[1] read in place of a missing row, zero selected
            const bool hb = 2 * q + 1 < nyt;
            v.a = p[0];
            const cpx<T> b = p[hb ? 1 : 0];
            v.b = {hb ? b.x : T(0), hb ? b.y : T(0)};
            return v;
        },
        [&](int e, Pair2<cpx<T>> v) {
            const int k = e / npair, q = e - k * npair;
            const cpx<T> a = v.a, b = v.b;
            s[q * P + k] = {a.x - b.y, -(a.y + b.x)};
            if (k > 0 && P - k >= H) s[q * P + P - k] = {a.x + b.y, a.y - b.x};
        });
    __syncthreads();
    s = fft_rows<T>(s, s1, npair, 1, P, in.plx, tw);
    for (int e = threadIdx.x; e < nyt * Lx; e += kThreads) {
        const int yy = e / Lx, x = e - yy * Lx;
        const cpx<T> v = s[(yy >> 1) * P + x];
        dst[(long)(y0 + yy) * Lx + x] = (yy & 1) ? -v.y : v.x;
    }
}

// Fixed-plan x passes (magnetisation mode only): Px == N for every item, a
// tile of 2*ROWS real rows packed into ROWS complex rows, transformed in
// place in one LDS buffer (rows past the tile's last pair are zero).
template <typename T, int N, int ROWS>
__global__ __launch_bounds__(kThreads) void k_xfwd_n(const Inst *__restrict__ I,
                                                     const Work *__restrict__ W,
                                                     const T *__restrict__ realb,
                                                     cpx<T> *__restrict__ specb,
                                                     const cpx<T> *__restrict__ twb)
{
    cpx<T> *s = reinterpret_cast<cpx<T> *>(g_smem);
    cpx<T> *tw = s + ROWS * N;
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const bool p16 = pair16<T>(in);
    const int c = w.a, z = w.b, y0 = w.c;
    constexpr int H = N / 2 + 1;
    const int Lx = in.nx, Ly = in.ny, Lz = in.nz;
    const T *src = realb + in.m_off + (long)c * in.N + (long)z * Lx * Ly;
    cpx<T> *dst = specb + in.a_off + (long)c * H * Lz * Ly;
    const int nyt = min(2 * ROWS, Ly - y0);
    const auto twr = load_twiddles<T, N>(twb + in.twx);
    staged_copy<ROWS * N, kThreads, cpx<T>>(
        ROWS * N,
        [&](int e) {
            const int q = e / N, x = e % N;
            T re = T(0), im = T(0);
            if (x < Lx && 2 * q < nyt) {
                const int ya = y0 + 2 * q;
                re = src[(long)ya * Lx + x];
                if (2 * q + 1 < nyt) im = src[(long)(ya + 1) * Lx + x];
            }
            return cpx<T>{re, im};
        },
        [&](int e, cpx<T> v) { s[e] = v; });
    store_twiddles<T, N>(tw, twr);
    __syncthreads();
    fft_fixed<T, N, ROWS>(s, tw);
    const int npair = (nyt + 1) >> 1;
    for (int e = threadIdx.x; e < H * npair; e += kThreads) {
        const int k = e / npair, q = e - k * npair;
        const cpx<T> A = s[q * N + k];
        const cpx<T> B = s[q * N + (k ? N - k : 0)];
        cpx<T> *o = dst + ((long)k * Lz + z) * Ly + y0 + 2 * q;
        const cpx<T> oa = {(A.x + B.x) * T(0.5), (A.y - B.y) * T(0.5)};
        const cpx<T> ob = {(A.y + B.y) * T(0.5), (B.x - A.x) * T(0.5)};
        if (p16) {
            store_pair16(o, oa, ob);
        } else {
            o[0] = oa;
            if (2 * q + 1 < nyt) o[1] = ob;
        }
    }
}

template <typename T, int N, int ROWS>
__global__ __launch_bounds__(kThreads) void k_xinv_n(const Inst *__restrict__ I,
                                                     const Work *__restrict__ W,
                                                     const cpx<T> *__restrict__ specb,
                                                     T *__restrict__ realb,
                                                     const cpx<T> *__restrict__ twb)
{
    cpx<T> *s = reinterpret_cast<cpx<T> *>(g_smem);
    cpx<T> *tw = s + ROWS * N;
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const bool p16 = pair16<T>(in);
    const int c = w.a, z = w.b, y0 = w.c;
    constexpr int H = N / 2 + 1;
    const int Lx = in.nx, Ly = in.ny, Lz = in.nz;
    const cpx<T> *src = specb + in.a_off + (long)c * H * Lz * Ly;
    T *dst = realb + in.h_off + (long)c * in.N + (long)z * Lx * Ly;
    const int nyt = min(2 * ROWS, Ly - y0);
    const int npair = (nyt + 1) >> 1;
    const auto twr = load_twiddles<T, N>(twb + in.twx);
    for (int e = threadIdx.x; e < (ROWS - npair) * N; e += kThreads) s[npair * N + e] = {T(0), T(0)};
    staged_copy<H * ROWS, kThreads, Pair2<cpx<T>>>(
        H * npair,
        [&](int e) {
            const int k = e / npair, q = e - k * npair;
            const cpx<T> *p = src + ((long)k * Lz + z) * Ly + y0 + 2 * q;
            if (p16) return load_pair16(p);
            Pair2<cpx<T>> v;
            // branch-free: p[1] read in place of a missing row, zero selected
            const bool hb = 2 * q + 1 < nyt;
            v.a = p[0];
            const cpx<T> b = p[hb ? 1 : 0];
            v.b = {hb ? b.x : T(0), hb ? b.y : T(0)};
            return v;
        },
        [&](int e, Pair2<cpx<T>> v) {
            const int k = e / npair, q = e - k * npair;
            const cpx<T> a = v.a, b = v.b;
            s[q * N + k] = {a.x - b.y, -(a.y + b.x)};
            if (k > 0 && N - k >= H) s[q * N + N - k] = {a.x + b.y, a.y - b.x};
        });
    store_twiddles<T, N>(tw, twr);
    __syncthreads();
    fft_fixed<T, N, ROWS>(s, tw);
    for (int e = threadIdx.x; e < nyt * Lx; e += kThreads) {
        const int yy = e / Lx, x = e - yy * Lx;
        const cpx<T> v = s[(yy >> 1) * N + x];
        dst[(long)(y0 + yy) * Lx + x] = (yy & 1) ? -v.y : v.x;
    }
}

template <typename T>
hipError_t launch_xfwd(hipStream_t st, const Inst *I, const Work *W, int n, int mode, int fixed_n,
                       int rows, const T *real, cpx<T> *spec, const cpx<T> *tw, size_t lds)
{
    if (n <= 0) return hipSuccess;
    const dim3 g(n), b(kThreads);
    constexpr int RM = kXRows<T>::value;
#define MMC_XF_CASE(TT, NN, RR) \
    if (fixed_n == NN && rows == RR) { MMC_LAUNCH((k_xfwd_n<TT, NN, RR>), g, b, lds, st, I, W, real, spec, tw); return post(); }
    if (mode == 0) { MMC_XFIXED(MMC_XF_CASE, T, 1) MMC_XFIXED(MMC_XF_CASE, T, RM) }
#undef MMC_XF_CASE
    MMC_LAUNCH(k_xfwd<T>, g, b, lds, st, I, W, mode, real, spec, tw);
    return post();
}

template <typename T>
hipError_t launch_xinv(hipStream_t st, const Inst *I, const Work *W, int n, int fixed_n, int rows,
                       const cpx<T> *spec, T *real, const cpx<T> *tw, size_t lds)
{
    if (n <= 0) return hipSuccess;
    const dim3 g(n), b(kThreads);
    constexpr int RM = kXRows<T>::value;
#define MMC_XI_CASE(TT, NN, RR) \
    if (fixed_n == NN && rows == RR) { MMC_LAUNCH((k_xinv_n<TT, NN, RR>), g, b, lds, st, I, W, spec, real, tw); return post(); }
    MMC_XFIXED(MMC_XI_CASE, T, 1) MMC_XFIXED(MMC_XI_CASE, T, RM)
#undef MMC_XI_CASE
    MMC_LAUNCH(k_xinv<T>, g, b, lds, st, I, W, spec, real, tw);
    return post();
}

template <typename T>
hipError_t set_lds_xpass(int bytes)
{
    hipError_t e = hipSuccess;
    set_lds_one(e, (const void *)k_xfwd<T>, bytes);
    set_lds_one(e, (const void *)k_xinv<T>, bytes);
    constexpr int RM = kXRows<T>::value;
#define MMC_XL(TT, NN, RR) set_lds_one(e, (const void *)k_xfwd_n<TT, NN, RR>, bytes); \
    set_lds_one(e, (const void *)k_xinv_n<TT, NN, RR>, bytes);
    MMC_XFIXED(MMC_XL, T, 1) MMC_XFIXED(MMC_XL, T, RM)
#undef MMC_XL
    return e;
}

#define MMC_INST(T)                                                                           \
    template hipError_t launch_xfwd<T>(hipStream_t, const Inst *, const Work *, int, int, int, \
                                       int, const T *, cpx<T> *, const cpx<T> *, size_t);     \
    template hipError_t launch_xinv<T>(hipStream_t, const Inst *, const Work *, int, int, int, \
                                       const cpx<T> *, T *, const cpx<T> *, size_t);          \
    template hipError_t set_lds_xpass<T>(int);
MMC_INST(float)
MMC_INST(double)
#undef MMC_INST

}  // namespace mmc
