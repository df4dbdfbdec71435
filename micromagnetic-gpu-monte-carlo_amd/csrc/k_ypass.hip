// k_ypass.hip -- z transforms of 3-D grids, the y pass (y transform,
// symmetric 3x3 k-space product of src/demag_cu.cu:68-116, inverse y
// transform), the y pass with the z convolution summed directly for films
// of a few layers, and the tensor's y transform.
#include "mmc_device.h"

namespace mmc {

// z transform: item (inst, c, kx, y0).  Short z (Pz <= kZDirect): one thread
// per y column, direct DFT in registers, item covers kThreads y values.
// Otherwise LDS Stockham with element (y, z) at s[z*nyt + y].
template <typename T>
__global__ __launch_bounds__(kThreads) void k_zfwd(const Inst *__restrict__ I,
                                                   const Work *__restrict__ W, int mode,
                                                   const cpx<T> *__restrict__ srcb,
                                                   cpx<T> *__restrict__ dstb,
                                                   const cpx<T> *__restrict__ twb)
{
    cpx<T> *s = reinterpret_cast<cpx<T> *>(g_smem);
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const int c = w.a, kx = w.b, y0 = w.c;
    const int H = in.Hx, Pz = in.Pz;
    int Ly, Lz, ty;
    const cpx<T> *src;
    cpx<T> *dst;
    if (mode == 0) {
        Ly = in.ny; Lz = in.nz; ty = in.tyz;
        src = srcb + in.a_off + ((long)c * H + kx) * Lz * Ly;
        dst = dstb + in.b_off + ((long)c * H + kx) * Pz * Ly;
    } else {
        Ly = in.Py; Lz = in.Pz; ty = in.tyz64;
        src = srcb + in.ka_off + ((long)c * H + kx) * Lz * Ly;
        dst = dstb + in.ka_off + ((long)c * H + kx) * Pz * Ly;
    }
    if (Pz <= kZDirect) {   // short z: direct DFT of one y column per thread
        const int y = y0 + threadIdx.x;
        if (y >= Ly) return;
        zdft_column<T, false>(src + y, dst + y, Ly, Lz, Pz, Pz, twb + in.twz);
        return;
    }
    cpx<T> *s1 = s + Pz * ty;
    const cpx<T> *tw = lds_twiddles(s1 + Pz * ty, twb + in.twz, Pz);
    const int nyt = min(ty, Ly - y0);
    staged_copy_dyn<cpx<T>>(
        Pz * nyt,
        [&](int e) {
            const int zz = e / nyt, yy = e - zz * nyt;
            T vx = T(0), vy = T(0);
            if (zz < Lz) {
                const cpx<T> t = src[(long)zz * Ly + y0 + yy];
                vx = t.x;
                vy = t.y;
            }
            return cpx<T>{vx, vy};
        },
        [&](int e, cpx<T> v) { s[e] = v; });
    __syncthreads();
    s = fft_rows<T>(s, s1, nyt, nyt, 1, in.plz, tw);
    for (int e = threadIdx.x; e < Pz * nyt; e += kThreads) {
        const int zz = e / nyt, yy = e - zz * nyt;
        dst[(long)zz * Ly + y0 + yy] = s[e];
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_zinv(const Inst *__restrict__ I,
                                                   const Work *__restrict__ W,
                                                   const cpx<T> *__restrict__ srcb,
                                                   cpx<T> *__restrict__ dstb,
                                                   const cpx<T> *__restrict__ twb)
{
    cpx<T> *s = reinterpret_cast<cpx<T> *>(g_smem);
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const int c = w.a, kx = w.b, y0 = w.c;
    const int H = in.Hx, Pz = in.Pz, Ly = in.ny, Lz = in.nz;
    cpx<T> *s1 = s + Pz * in.tyz;
    const cpx<T> *src = srcb + in.b_off + ((long)c * H + kx) * Pz * Ly;
    cpx<T> *dst = dstb + in.a_off + ((long)c * H + kx) * Lz * Ly;
    if (Pz <= kZDirect) {
        const int y = y0 + threadIdx.x;
        if (y >= Ly) return;
        zdft_column<T, true>(src + y, dst + y, Ly, Pz, Lz, Pz, twb + in.twz);
        return;
    }
    const cpx<T> *tw = lds_twiddles(s1 + Pz * in.tyz, twb + in.twz, Pz);
    const int nyt = min(in.tyz, Ly - y0);
    staged_copy_dyn<cpx<T>>(
        Pz * nyt,
        [&](int e) {
            const int zz = e / nyt, yy = e - zz * nyt;
            return src[(long)zz * Ly + y0 + yy];
        },
        [&](int e, cpx<T> v) { s[e] = cconj(v); });
    __syncthreads();
    s = fft_rows<T>(s, s1, nyt, nyt, 1, in.plz, tw);
    for (int e = threadIdx.x; e < Lz * nyt; e += kThreads) {
        const int zz = e / nyt, yy = e - zz * nyt;
        dst[(long)zz * Ly + y0 + yy] = cconj(s[e]);
    }
}

// Compile-time-planned z transforms (Pz = PZ, a fixed_z_size, magnetisation
// only): item (inst, c, kx, y0) covers TY = z_cols y columns; the tile is [PZ][TY]
// in LDS (z outer, y contiguous), loaded and stored with y along the lanes,
// transformed along z with neighbouring lanes on neighbouring columns.
// Same arithmetic as the runtime-plan path (same radix order and twiddles).
// HALF: every item's instance has 2 nz <= PZ (an open z axis), so the
// forward transform reads only the lower half of the z rows (the upper half
// is padding) and the inverse forms only the nz rows it stores.
template <typename T, int PZ, bool HALF>
__global__ __launch_bounds__(kThreads) void k_zfwd_n(const Inst *__restrict__ I,
                                                     const Work *__restrict__ W,
                                                     const cpx<T> *__restrict__ srcb,
                                                     cpx<T> *__restrict__ dstb,
                                                     const cpx<T> *__restrict__ twb)
{
    constexpr int TY = z_cols(sizeof(T));
    cpx<T> *s = reinterpret_cast<cpx<T> *>(g_smem);
    cpx<T> *tw = s + PZ * TY;
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const int c = w.a, kx = w.b, y0 = w.c;
    const int H = in.Hx, Ly = in.ny, Lz = in.nz;
    const cpx<T> *src = srcb + in.a_off + ((long)c * H + kx) * Lz * Ly + y0;
    cpx<T> *dst = dstb + in.b_off + ((long)c * H + kx) * PZ * Ly + y0;
    const int nyt = min(TY, Ly - y0);
    const auto twr = load_twiddles<T, PZ>(twb + in.twz);
    constexpr int ZL = HALF ? PZ / 2 : PZ;   // z rows staged
    staged_copy<ZL * TY, kThreads, cpx<T>>(
        ZL * TY,
        [&](int e) {
            const int z = e / TY, y = e - z * TY;
            T vx = T(0), vy = T(0);
            if (z < Lz && y < nyt) {
                const cpx<T> t = src[(long)z * Ly + y];
                vx = t.x;
                vy = t.y;
            }
            return cpx<T>{vx, vy};
        },
        [&](int e, cpx<T> v) { s[e] = v; });
    store_twiddles<T, PZ>(tw, twr);
    __syncthreads();
    fft_fixed_cols<T, PZ, TY, kThreads, HALF>(s, tw);
    for (int e = threadIdx.x; e < PZ * TY; e += kThreads) {
        const int z = e / TY, y = e - z * TY;
        if (y < nyt) dst[(long)z * Ly + y] = s[e];
    }
}

template <typename T, int PZ, bool HALF>
__global__ __launch_bounds__(kThreads) void k_zinv_n(const Inst *__restrict__ I,
                                                     const Work *__restrict__ W,
                                                     const cpx<T> *__restrict__ srcb,
                                                     cpx<T> *__restrict__ dstb,
                                                     const cpx<T> *__restrict__ twb)
{
    constexpr int TY = z_cols(sizeof(T));
    cpx<T> *s = reinterpret_cast<cpx<T> *>(g_smem);
    cpx<T> *tw = s + PZ * TY;
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const int c = w.a, kx = w.b, y0 = w.c;
    const int H = in.Hx, Ly = in.ny, Lz = in.nz;
    const cpx<T> *src = srcb + in.b_off + ((long)c * H + kx) * PZ * Ly + y0;
    cpx<T> *dst = dstb + in.a_off + ((long)c * H + kx) * Lz * Ly + y0;
    const int nyt = min(TY, Ly - y0);
    const auto twr = load_twiddles<T, PZ>(twb + in.twz);
    staged_copy<PZ * TY, kThreads, cpx<T>>(
        PZ * TY,
        [&](int e) {
            const int z = e / TY, y = e - z * TY;
            T vx = T(0), vy = T(0);
            if (y < nyt) {
                const cpx<T> t = src[(long)z * Ly + y];
                vx = t.x;
                vy = t.y;
            }
            return cpx<T>{vx, vy};
        },
        [&](int e, cpx<T> v) { s[e] = cconj(v); });
    store_twiddles<T, PZ>(tw, twr);
    __syncthreads();
    fft_fixed_cols<T, PZ, TY, kThreads, false, HALF && fixed_olo_ok(PZ)>(s, tw);
    for (int e = threadIdx.x; e < Lz * TY; e += kThreads) {
        const int z = e / TY, y = e - z * TY;
        if (y < nyt) dst[(long)z * Ly + y] = cconj(s[e]);
    }
}

// Khat component c6 at k point ki of instance `in`.  The instance's Khat
// slot holds six planes of S = Hx*Pz*Py complex values; a plane stored real
// (kmode 1) uses its first S scalars, a zero plane (kmode 0) is not read.
template <typename T>
__device__ __forceinline__ cpx<T> khat_at(const T *__restrict__ kr, const Inst &in, int c6,
                                          long S, long ki)
{
    const int md = in.kmode[c6];
    const T *p = kr + 2 * (in.khat_off + c6 * S);
    if (md == 2) return reinterpret_cast<const cpx<T> *>(p)[ki];
    if (md == 1) return {p[ki], T(0)};
    return {T(0), T(0)};
}

// khat_at split in two for values loaded early and used late: khat_raw
// issues one 8-byte (FP64: 16-byte) load from an address chosen by the
// plane's storage (complex: the value; real: the aligned pair of scalars
// holding it; zero: the plane's first element, the same for every lane), and
// khat_val forms the value where it is used.  No load is selected or
// branched around, so no load waits for another.
template <typename T>
__device__ __forceinline__ cpx<T> khat_raw(const T *__restrict__ kr, const Inst &in, int c6,
                                           long S, long ki)
{
    const int md = in.kmode[c6];
    const cpx<T> *p = reinterpret_cast<const cpx<T> *>(kr + 2 * (in.khat_off + c6 * S));
    return p[md == 2 ? ki : (md == 1 ? (ki >> 1) : 0)];
}

template <typename T>
__device__ __forceinline__ cpx<T> khat_val(const cpx<T> &raw, const Inst &in, int c6, long ki)
{
    const int md = in.kmode[c6];
    if (md == 2) return raw;
    return {md == 1 ? ((ki & 1) ? raw.y : raw.x) : T(0), T(0)};
}

// Symmetric 3x3 complex product (src/demag_cu.cu:98-111) of one k point,
// in place, conjugated for the inverse transform that follows.
template <typename T>
__device__ __forceinline__ void kmul_conj(cpx<T> *px, cpx<T> *py, cpx<T> *pz, const cpx<T> *k6)
{
    const cpx<T> mx = *px, my = *py, mz = *pz;
    const cpx<T> hx = cdot3(mx, k6[0], my, k6[1], mz, k6[2]);
    const cpx<T> hy = cdot3(mx, k6[1], my, k6[3], mz, k6[4]);
    const cpx<T> hz = cdot3(mx, k6[2], my, k6[4], mz, k6[5]);
    *px = cconj(hx);
    *py = cconj(hy);
    *pz = cconj(hz);
}

// Compile-time-planned y pass for Py = N over up to C columns: one LDS
// buffer of 3*C rows (row comp*C + col), the item's Khat values prefetched
// into registers before the forward transform so their HBM latency overlaps
// it.
template <typename T, int N, int C, bool HALF>
__device__ __forceinline__ void ymul_fixed(const Inst &in, const Work &w, cpx<T> *__restrict__ bufb,
                                           const cpx<T> *__restrict__ khatb,
                                           const cpx<T> *__restrict__ twb, int a2)
{
    // HALF: the instance's ny rows fit in the lower half of the padded
    // length (an open y axis, N = 2ny), so the forward transform reads only
    // that half (its upper half is the zero padding) and the inverse forms
    // only that half
    constexpr bool OLO = HALF && fixed_olo_ok(N);
    constexpr int NL = HALF ? N / 2 : N;             // rows' elements loaded per column
    constexpr int ROWS = 3 * C;
    constexpr int Q = (C * N + kThreads - 1) / kThreads;
    // the k points of a thread are prefetched across the forward transform
    // when their six values fit in 24 VGPRs (two FP32 / one FP64 point); more
    // would cost the registers that keep several workgroups resident per
    // SIMD, so they are loaded together right after it
    constexpr bool PF = Q * 6 * (int)sizeof(cpx<T>) <= 96;
    cpx<T> *s = reinterpret_cast<cpx<T> *>(g_smem);
    cpx<T> *tw = s + ROWS * N;
    const int kx0 = w.a, kz = w.b, ncol = w.c;
    const int H = in.Hx, Pz = in.Pz, Ly = in.ny;
    cpx<T> *buf = bufb + in.b_off;
    const T *kr = reinterpret_cast<const T *>(khatb);
    const long cstride = (long)H * Pz * Ly;
    const long kstride = (long)H * Pz * N;
    // k points past the item's columns read the last column's values (a
    // valid address; the product of such lanes is never stored), so the
    // loads need no select, which would wait for each of them in turn
    auto load_k6 = [&](int q, cpx<T> *k6) {
        const int e = threadIdx.x + q * kThreads;
        const int col = min(e / N, ncol - 1), ky = e % N;
        const long ki = ((long)(kx0 + col) * Pz + kz) * N + ky;
        #pragma unroll
        for (int c6 = 0; c6 < 6; ++c6) k6[c6] = khat_raw(kr, in, c6, kstride, ki);
    };
    // loads in the order their data is needed (vector loads complete in
    // issue order): twiddles and spectrum rows, then the tensor values,
    // which may still be in flight during the forward transform
    const auto twr = load_twiddles<T, N>(twb + in.twy);
    const auto rows = staged_load<ROWS * NL, kThreads, cpx<T>>(ROWS * NL, [&](int e) {
        const int r = e / NL, y = e % NL;
        const int comp = r / C, col = r % C;
        // component-wise select: a conditionally loaded struct is kept in scratch
        T vx = T(0), vy = T(0);
        if (col < ncol && y < Ly) {
            const cpx<T> t = buf[comp * cstride + ((long)(kx0 + col) * Pz + kz) * Ly + y];
            vx = t.x;
            vy = t.y;
        }
        return cpx<T>{vx, vy};
    });
    cpx<T> k6p[PF ? Q : 1][6];
    if (PF) {
        #pragma unroll
        for (int q = 0; q < (PF ? Q : 1); ++q) load_k6(q, k6p[q]);
    }
    staged_store<ROWS * NL, kThreads, cpx<T>>(ROWS * NL, rows, [&](int e, cpx<T> v) {
        s[(e / NL) * N + e % NL] = v;
    });
    store_twiddles<T, N>(tw, twr);
    __syncthreads();
    fft_fixed<T, N, ROWS, kThreads, false, HALF>(s, tw);
    // without prefetch, all k points' values are loaded before the first product
    cpx<T> k6l[PF ? 1 : Q][6];
    if (!PF) {
        #pragma unroll
        for (int q = 0; q < (PF ? 1 : Q); ++q) load_k6(q, k6l[q]);
    }
    #pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = threadIdx.x + q * kThreads;
        if (e < C * N) {
            const int col = e / N, ky = e % N;
            cpx<T> k6[6];
            if (PF) {
                #pragma unroll
                for (int c6 = 0; c6 < 6; ++c6) k6[c6] = khat_val(k6p[PF ? q : 0][c6], in, c6, ky);
            } else {
                #pragma unroll
                for (int c6 = 0; c6 < 6; ++c6) k6[c6] = khat_val(k6l[PF ? 0 : q][c6], in, c6, ky);
            }
            kmul_conj(s + (0 * C + col) * N + ky, s + (1 * C + col) * N + ky,
                      s + (2 * C + col) * N + ky, k6);
        }
    }
    __syncthreads();
    fft_fixed<T, N, ROWS, kThreads, false, false, OLO>(s, tw);
    if (a2 && in.a2) {
        // pair-major, for the inverse x pass of a fused sweep (Pz == 1):
        // element (c, y) at P[((c*ny/2 + y/2)*Hx + kx)*2 + (y & 1)], the
        // item's columns side by side
        cpx<T> *P = bufb + in.pm_off;
        const int np = Ly >> 1;
        for (int e = threadIdx.x; e < 3 * np * C; e += kThreads) {
            const int col = e % C, rp = e / C;
            const int comp = rp / np, p = rp - comp * np;
            if (col >= ncol) continue;
            const int r = comp * C + col;
            const cpx<T> a = cconj(s[r * N + 2 * p]), b = cconj(s[r * N + 2 * p + 1]);
            cpx<T> *o = P + (((long)comp * np + p) * H + kx0 + col) * 2;
            if (sizeof(T) == 4) {
                store_pair16(o, a, b);
            } else {
                o[0] = a;
                o[1] = b;
            }
        }
        return;
    }
    const Div dLy(Ly);
    for (int e = threadIdx.x; e < ROWS * Ly; e += kThreads) {
        const int r = dLy(e), y = e - r * Ly;
        const int comp = r / C, col = r % C;
        if (col < ncol)
            buf[comp * cstride + ((long)(kx0 + col) * Pz + kz) * Ly + y] = cconj(s[r * N + y]);
    }
}

// Fixed-plan y pass: every item's instance has Py == N and at most C
// columns (one launch per distinct (length, C); a kernel of its own keeps its
// register count low).  C is fixed_cols(N), or 1 for grids too small to fill
// the chip with wide items.
template <typename T, int N, int C, bool HALF>
__global__ __launch_bounds__(kThreads) void k_ymul_n(const Inst *__restrict__ I,
                                                     const Work *__restrict__ W,
                                                     cpx<T> *__restrict__ bufb,
                                                     const cpx<T> *__restrict__ khatb,
                                                     const cpx<T> *__restrict__ twb, int a2)
{
    const Work w = W[blockIdx.x];
    ymul_fixed<T, N, C, HALF>(I[w.inst], w, bufb, khatb, twb, a2);
}

// One launch over y-pass items of every y length of a batch (any_y_len),
// columns per item from tile policy POL (any_tile).
template <typename T, int POL, bool HALF>
__global__ __launch_bounds__(kThreads) void k_ymul_any(const Inst *__restrict__ I,
                                                       const Work *__restrict__ W,
                                                       cpx<T> *__restrict__ bufb,
                                                       const cpx<T> *__restrict__ khatb,
                                                       const cpx<T> *__restrict__ twb, int a2)
{
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    switch (in.Py) {
#define MMC_ANY(NN) case NN: ymul_fixed<T, NN, any_tile(POL, NN), HALF>(in, w, bufb, khatb, twb, a2); break;
    MMC_ANY(64) MMC_ANY(96) MMC_ANY(128) MMC_ANY(192) MMC_ANY(256) MMC_ANY(384) MMC_ANY(512)
#undef MMC_ANY
    }
}

// y forward transform, k-space product with a direct z convolution, y
// inverse of one x frequency of an instance with a few layers (nz = NZ <=
// kZConvMax, Pz <= kZDirect, Py = N).  The tensor's z slices are transformed
// along x and y only, slot t of Khat holding slice offset t (mod Pz), so
//   H^(kx, ky, z) = sum_{z' < nz} Khat[(z - z') mod Pz](kx, ky) M^(kx, ky, z')
// for the nz real layers: the cyclic convolution along z that a z transform
// of length Pz would compute, summed directly, with no z transform and no
// xz spectrum.  Item (inst, kx, 0, 0); reads and writes the x spectrum A;
// LDS row (z, comp) = z*3 + comp.
// Waves per SIMD the compiler keeps room for in k_yzconv_n: six (three
// workgroups per CU, 80 VGPRs) for FP32 films of two layers up to 1024
// points, whose LDS allows three; the LDS of larger variants allows fewer
// anyway.
template <typename T, int N, int NZ>
constexpr int yzconv_waves() { return sizeof(T) == 4 && NZ == 2 && N <= 1024 ? 6 : 1; }

template <typename T, int N, int NZ>
__global__ __launch_bounds__(kYzThreads) __attribute__((amdgpu_waves_per_eu(yzconv_waves<T, N, NZ>())))
void k_yzconv_n(const Inst *__restrict__ I,
                                                       const Work *__restrict__ W,
                                                       cpx<T> *__restrict__ Ab,
                                                       const cpx<T> *__restrict__ khatb,
                                                       const cpx<T> *__restrict__ twb, int a2)
{
    constexpr int NT = kYzThreads;
    constexpr int ROWS = 3 * NZ;
    constexpr int ND = 2 * NZ - 1;                 // z offsets -(NZ-1) .. NZ-1
    constexpr int Q = (N + NT - 1) / NT;           // k points per thread
    // The rows and a quarter twiddle table (50 KB for two FP32 layers of
    // 1024) leave room for three workgroups per CU, so a 1024-point x axis
    // (513 items) runs in one round; the k points' tensor values are
    // prefetched across the forward transform only when that costs no
    // occupancy (24 VGPRs).
    constexpr bool PF = Q == 1 && ND * 6 * (int)sizeof(cpx<T>) <= 96;
    cpx<T> *s = reinterpret_cast<cpx<T> *>(g_smem);
    cpx<T> *tw = s + ROWS * N;                     // tw[u] = exp(-2 pi i u / N), u < N/4
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const int kx = w.a;
    const int Pz = in.Pz, Ly = in.ny, H = in.Hx;
    cpx<T> *A = Ab + in.a_off;
    const long astride = (long)H * NZ * Ly;        // component stride in A
    const T *kr = reinterpret_cast<const T *>(khatb);
    const long kstride = (long)H * Pz * N;         // component stride in Khat
    // lanes past the length read the last k point (a valid address; their
    // products are never formed), so the loads need no select
    auto load_k = [&](int q, cpx<T> (&k6)[ND][6]) {
        const int ky = min(threadIdx.x + q * NT, N - 1);
        #pragma unroll
        for (int d = 0; d < ND; ++d) {
            const int slot = (d - (NZ - 1) + Pz) % Pz;
            const long ki = ((long)kx * Pz + slot) * N + ky;
            #pragma unroll
            for (int c6 = 0; c6 < 6; ++c6) k6[d][c6] = khat_raw(kr, in, c6, kstride, ki);
        }
    };
    const auto twr = load_twiddles<T, N / 4, NT>(twb + in.twy);
    cpx<T> kp[PF ? Q : 1][ND][6];
    auto row_ld = [&](int e) {
        const int r = e / N, y = e - r * N;
        const int z = r / 3, c = r - 3 * z;
        T vx = T(0), vy = T(0);
        if (y < Ly) {
            const cpx<T> t = A[c * astride + ((long)kx * NZ + z) * Ly + y];
            vx = t.x;
            vy = t.y;
        }
        return cpx<T>{vx, vy};
    };
    auto row_st = [&](int e, cpx<T> v) { s[e] = v; };
    if (PF) {
        const auto rows = staged_load<ROWS * N, NT, cpx<T>>(ROWS * N, row_ld);
        load_k(0, kp[0]);
        staged_store<ROWS * N, NT, cpx<T>>(ROWS * N, rows, row_st);
    } else {
        staged_copy<ROWS * N, NT, cpx<T>>(ROWS * N, row_ld, row_st);
    }
    store_twiddles<T, N / 4, NT>(tw, twr);
    __syncthreads();
    fft_fixed<T, N, ROWS, NT, true>(s, tw);
    // one k point at a time (a rolled loop keeps the next point's tensor
    // loads from being hoisted, which would cost the third workgroup)
    #pragma unroll 1
    for (int q = 0; q < Q; ++q) {
        const int ky = threadIdx.x + q * NT;
        if (ky >= N) break;
        cpx<T> kq[ND][6];
        if (PF) {
            #pragma unroll
            for (int d = 0; d < ND; ++d)
                #pragma unroll
                for (int c6 = 0; c6 < 6; ++c6) kq[d][c6] = kp[0][d][c6];
        } else {
            load_k(q, kq);
        }
        #pragma unroll
        for (int d = 0; d < ND; ++d)
            #pragma unroll
            for (int c6 = 0; c6 < 6; ++c6) kq[d][c6] = khat_val(kq[d][c6], in, c6, ky);
        cpx<T> mv[NZ][3];
        #pragma unroll
        for (int z = 0; z < NZ; ++z)
            #pragma unroll
            for (int c = 0; c < 3; ++c) mv[z][c] = s[(z * 3 + c) * N + ky];
        #pragma unroll
        for (int z = 0; z < NZ; ++z) {
            cpx<T> h[3];
            #pragma unroll
            for (int zp = 0; zp < NZ; ++zp) {
                const cpx<T> *k6 = kq[z - zp + NZ - 1];
                const cpx<T> mx = mv[zp][0], my = mv[zp][1], mz = mv[zp][2];
                if (zp == 0) {
                    h[0] = cdot3(mx, k6[0], my, k6[1], mz, k6[2]);
                    h[1] = cdot3(mx, k6[1], my, k6[3], mz, k6[4]);
                    h[2] = cdot3(mx, k6[2], my, k6[4], mz, k6[5]);
                } else {
                    h[0] = cdot3_acc(h[0], mx, k6[0], my, k6[1], mz, k6[2]);
                    h[1] = cdot3_acc(h[1], mx, k6[1], my, k6[3], mz, k6[4]);
                    h[2] = cdot3_acc(h[2], mx, k6[2], my, k6[4], mz, k6[5]);
                }
            }
            #pragma unroll
            for (int c = 0; c < 3; ++c) s[(z * 3 + c) * N + ky] = cconj(h[c]);
        }
    }
    __syncthreads();
    fft_fixed<T, N, ROWS, NT, true>(s, tw);
    if (a2 && in.a2) {
        // pair-major, for the inverse x pass of a fused sweep: element
        // (c, z, y) at B[(((c*NZ + z)*ny/2 + y/2)*Hx + kx)*2 + (y & 1)], so a
        // row pair's spectrum is contiguous; the items of 8 consecutive x
        // frequencies run on one XCD and fill each 128-byte line together
        cpx<T> *A2 = Ab + in.pm_off;
        const int np = Ly >> 1;
        const Div dnp(np);
        for (int e = threadIdx.x; e < ROWS * np; e += NT) {
            const int r = dnp(e), p = e - r * np;
            const int z = r / 3, c = r - 3 * z;
            const cpx<T> a = cconj(s[r * N + 2 * p]), b = cconj(s[r * N + 2 * p + 1]);
            cpx<T> *o = A2 + ((((long)c * NZ + z) * np + p) * H + kx) * 2;
            if (sizeof(T) == 4) {
                store_pair16(o, a, b);
            } else {
                o[0] = a;
                o[1] = b;
            }
        }
        return;
    }
    const Div dLy(Ly);
    for (int e = threadIdx.x; e < ROWS * Ly; e += NT) {
        const int r = dLy(e), y = e - r * Ly;
        const int z = r / 3, c = r - 3 * z;
        A[c * astride + ((long)kx * NZ + z) * Ly + y] = cconj(s[r * N + y]);
    }
}

// y forward, symmetric 3x3 product with Khat, y inverse.  Item (inst, kx0,
// kz, ncol): columns kx0..kx0+ncol-1; generic ping-pong path with LDS row
// (comp, col) = comp*ncol+col.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_ymul(const Inst *__restrict__ I,
                                                   const Work *__restrict__ W,
                                                   cpx<T> *__restrict__ bufb,
                                                   const cpx<T> *__restrict__ khatb,
                                                   const cpx<T> *__restrict__ twb)
{
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    cpx<T> *s = reinterpret_cast<cpx<T> *>(g_smem);
    const int kx0 = w.a, kz = w.b, ncol = w.c;
    cpx<T> *s1 = s + 3 * in.ncol * in.Py;
    const cpx<T> *tw = lds_twiddles(s1 + 3 * in.ncol * in.Py, twb + in.twy, in.Py);
    const int H = in.Hx, Pz = in.Pz, Py = in.Py, Ly = in.ny;
    cpx<T> *buf = bufb + in.b_off;
    const T *kr = reinterpret_cast<const T *>(khatb);
    const long cstride = (long)H * Pz * Ly;    // component stride in B
    const long kstride = (long)H * Pz * Py;    // component stride in Khat
    const int rows = 3 * ncol;
    staged_copy_dyn<cpx<T>>(
        rows * Py,
        [&](int e) {
            const int r = e / Py, y = e - r * Py;
            const int comp = r / ncol, col = r - comp * ncol;
            T vx = T(0), vy = T(0);
            if (y < Ly) {
                const cpx<T> t = buf[comp * cstride + ((long)(kx0 + col) * Pz + kz) * Ly + y];
                vx = t.x;
                vy = t.y;
            }
            return cpx<T>{vx, vy};
        },
        [&](int e, cpx<T> v) { s[e] = v; });
    __syncthreads();
    s = fft_rows<T>(s, s1, rows, 1, Py, in.ply, tw);
    s1 = (s == s1) ? reinterpret_cast<cpx<T> *>(g_smem) : s1;
    for (int e = threadIdx.x; e < ncol * Py; e += kThreads) {
        const int col = e / Py, ky = e - col * Py;
        const long ki = ((long)(kx0 + col) * Pz + kz) * Py + ky;
        const cpx<T> kxx = khat_at(kr, in, 0, kstride, ki), kxy = khat_at(kr, in, 1, kstride, ki);
        const cpx<T> kxz = khat_at(kr, in, 2, kstride, ki), kyy = khat_at(kr, in, 3, kstride, ki);
        const cpx<T> kyz = khat_at(kr, in, 4, kstride, ki), kzz = khat_at(kr, in, 5, kstride, ki);
        cpx<T> *px = s + (0 * ncol + col) * Py + ky;
        cpx<T> *py = s + (1 * ncol + col) * Py + ky;
        cpx<T> *pz = s + (2 * ncol + col) * Py + ky;
        const cpx<T> mx = *px, my = *py, mz = *pz;
        const cpx<T> hx = cdot3(mx, kxx, my, kxy, mz, kxz);
        const cpx<T> hy = cdot3(mx, kxy, my, kyy, mz, kyz);
        const cpx<T> hz = cdot3(mx, kxz, my, kyz, mz, kzz);
        *px = cconj(hx);
        *py = cconj(hy);
        *pz = cconj(hz);
    }
    __syncthreads();
    s = fft_rows<T>(s, s1, rows, 1, Py, in.ply, tw);
    for (int e = threadIdx.x; e < rows * Ly; e += kThreads) {
        const int r = e / Ly, y = e - r * Ly;
        const int comp = r / ncol, col = r - comp * ncol;
        buf[comp * cstride + ((long)(kx0 + col) * Pz + kz) * Ly + y] = cconj(s[r * Py + y]);
    }
}

// Tensor y transform (FP64), item (inst, comp, kx, kz); writes Khat / NP
// (Khat / (Px Py) for zd instances, whose slices skip the z transform).
template <typename TO>
__global__ __launch_bounds__(kThreads) void k_yfwd_tensor(const Inst *__restrict__ I,
                                                          const Work *__restrict__ W,
                                                          const cpx<double> *__restrict__ kab,
                                                          cpx<TO> *__restrict__ khatb,
                                                          const cpx<double> *__restrict__ twb)
{
    cpx<double> *s = reinterpret_cast<cpx<double> *>(g_smem);
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const int comp = w.a, kx = w.b, kz = w.c;
    const int H = in.Hx, Pz = in.Pz, Py = in.Py;
    const long off = ((long)comp * H + kx) * Pz * Py + (long)kz * Py;
    const cpx<double> *src = kab + in.ka_off + off;
    cpx<TO> *dst = khatb + in.khat_off + off;
    const cpx<double> *tw = lds_twiddles(s + 2 * Py, twb + in.twy, Py);
    for (int e = threadIdx.x; e < Py; e += kThreads) s[e] = src[e];
    __syncthreads();
    s = fft_rows<double>(s, s + Py, 1, 1, Py, in.ply, tw);
    // no z transform for zd instances: their z slices are convolved directly
    const double scale = 1.0 / (in.zd ? (double)in.Px * (double)in.Py : (double)in.NP);
    for (int e = threadIdx.x; e < Py; e += kThreads)
        dst[e] = {TO(s[e].x * scale), TO(s[e].y * scale)};
}

template <typename T>
hipError_t launch_zfwd(hipStream_t st, const Inst *I, const Work *W, int n, int mode,
                       int fixed_pz, const cpx<T> *src, cpx<T> *dst, const cpx<T> *tw, size_t lds,
                       bool half)
{
    if (n <= 0) return hipSuccess;
    if (mode == 0 && fixed_z_size(fixed_pz)) {
        const dim3 g(n), b(kThreads);
#define MMC_ZF(PZ)                                                                             \
    if (fixed_pz == PZ) {                                                                      \
        if (half) MMC_LAUNCH((k_zfwd_n<T, PZ, true>), g, b, lds, st, I, W, src, dst, tw);     \
        else MMC_LAUNCH((k_zfwd_n<T, PZ, false>), g, b, lds, st, I, W, src, dst, tw);         \
        return post();                                                                         \
    }
        MMC_ZF(16) MMC_ZF(32) MMC_ZF(64)
#undef MMC_ZF
    }
    MMC_LAUNCH(k_zfwd<T>, dim3(n), dim3(kThreads), lds, st, I, W, mode, src, dst, tw);
    return post();
}

template <typename T>
hipError_t launch_zinv(hipStream_t st, const Inst *I, const Work *W, int n, int fixed_pz,
                       const cpx<T> *src, cpx<T> *dst, const cpx<T> *tw, size_t lds, bool half)
{
    if (n <= 0) return hipSuccess;
    if (fixed_z_size(fixed_pz)) {
        const dim3 g(n), b(kThreads);
#define MMC_ZI(PZ)                                                                             \
    if (fixed_pz == PZ) {                                                                      \
        if (half) MMC_LAUNCH((k_zinv_n<T, PZ, true>), g, b, lds, st, I, W, src, dst, tw);     \
        else MMC_LAUNCH((k_zinv_n<T, PZ, false>), g, b, lds, st, I, W, src, dst, tw);         \
        return post();                                                                         \
    }
        MMC_ZI(16) MMC_ZI(32) MMC_ZI(64)
#undef MMC_ZI
    }
    MMC_LAUNCH(k_zinv<T>, dim3(n), dim3(kThreads), lds, st, I, W, src, dst, tw);
    return post();
}

template <typename T>
hipError_t launch_ymul(hipStream_t st, const Inst *I, const Work *W, int n, int fixed_n,
                       int cols, cpx<T> *buf, const cpx<T> *khat, const cpx<T> *tw, size_t lds,
                       bool half, bool a2_)
{
    const int a2 = a2_;
    if (n <= 0) return hipSuccess;
    const dim3 g(n), b(kThreads);
    if (fixed_n == kAnyLen) {
#define MMC_YA(P)                                                                              \
    if (half) MMC_LAUNCH((k_ymul_any<T, P, true>), g, b, lds, st, I, W, buf, khat, tw, a2);        \
    else MMC_LAUNCH((k_ymul_any<T, P, false>), g, b, lds, st, I, W, buf, khat, tw, a2);
        switch (cols) {
        case 0: MMC_YA(0) break;
        case 1: MMC_YA(1) break;
        case 2: MMC_YA(2) break;
        default: MMC_YA(3) break;
        }
#undef MMC_YA
        return post();
    }
#define MMC_YN(NN)                                                                              \
    if (fixed_n == NN) {                                                                        \
        if (half && cols == 1) MMC_LAUNCH((k_ymul_n<T, NN, 1, true>), g, b, lds, st, I, W, buf, khat, tw, a2); \
        else if (half) MMC_LAUNCH((k_ymul_n<T, NN, fixed_cols(NN), true>), g, b, lds, st, I, W, buf, khat, tw, a2); \
        else if (cols == 1) MMC_LAUNCH((k_ymul_n<T, NN, 1, false>), g, b, lds, st, I, W, buf, khat, tw, a2); \
        else MMC_LAUNCH((k_ymul_n<T, NN, fixed_cols(NN), false>), g, b, lds, st, I, W, buf, khat, tw, a2); \
        return post();                                                                          \
    }
    MMC_YN(64) MMC_YN(96) MMC_YN(128) MMC_YN(192) MMC_YN(256) MMC_YN(384) MMC_YN(512)
    MMC_YN(768) MMC_YN(1024) MMC_YN(2048)
#undef MMC_YN
    MMC_LAUNCH(k_ymul<T>, g, b, lds, st, I, W, buf, khat, tw);
    return post();
}

template <typename T>
hipError_t launch_yzconv(hipStream_t st, const Inst *I, const Work *W, int n, int fixed_n, int nz,
                         cpx<T> *A, const cpx<T> *khat, const cpx<T> *tw, size_t lds, bool a2)
{
    if (n <= 0) return hipSuccess;
    const dim3 g(n);
#define MMC_YZC(NN, ZZ)                                                                        \
    if (fixed_n == NN && nz == ZZ) {                                                           \
        MMC_LAUNCH((k_yzconv_n<T, NN, ZZ>), g, dim3(kYzThreads), lds, st, I, W, A, khat, tw,   \
                   (int)a2);                                                                   \
        return post();                                                                         \
    }
#define MMC_YZC_N(NN) MMC_YZC(NN, 2) MMC_YZC(NN, 3) MMC_YZC(NN, 4)
    MMC_YZC_N(64) MMC_YZC_N(96) MMC_YZC_N(128) MMC_YZC_N(192) MMC_YZC_N(256) MMC_YZC_N(384)
    MMC_YZC_N(512) MMC_YZC_N(768) MMC_YZC_N(1024) MMC_YZC_N(2048)
#undef MMC_YZC_N
#undef MMC_YZC
    return hipErrorInvalidValue;
}

template <typename TO>
hipError_t launch_yfwd_tensor(hipStream_t st, const Inst *I, const Work *W, int n,
                              const cpx<double> *ka, cpx<TO> *khat, const cpx<double> *tw,
                              size_t lds)
{
    if (n <= 0) return hipSuccess;
    MMC_LAUNCH(k_yfwd_tensor<TO>, dim3(n), dim3(kThreads), lds, st, I, W, ka, khat, tw);
    return post();
}

template <typename T>
hipError_t set_lds_ypass(int bytes)
{
    hipError_t e = hipSuccess;
    set_lds_one(e, (const void *)k_zfwd<T>, bytes);
    set_lds_one(e, (const void *)k_zinv<T>, bytes);
#define MMC_ZL(PZ) set_lds_one(e, (const void *)k_zfwd_n<T, PZ, false>, bytes); \
    set_lds_one(e, (const void *)k_zinv_n<T, PZ, false>, bytes);                  \
    set_lds_one(e, (const void *)k_zfwd_n<T, PZ, true>, bytes);                   \
    set_lds_one(e, (const void *)k_zinv_n<T, PZ, true>, bytes);
    MMC_ZL(16) MMC_ZL(32) MMC_ZL(64)
#undef MMC_ZL
    set_lds_one(e, (const void *)k_ymul<T>, bytes);
#define MMC_YL(NN) set_lds_one(e, (const void *)k_ymul_n<T, NN, 1, false>, bytes); \
    set_lds_one(e, (const void *)k_ymul_n<T, NN, fixed_cols(NN), false>, bytes);      \
    set_lds_one(e, (const void *)k_ymul_n<T, NN, 1, true>, bytes);                    \
    set_lds_one(e, (const void *)k_ymul_n<T, NN, fixed_cols(NN), true>, bytes);
    MMC_YL(64) MMC_YL(96) MMC_YL(128) MMC_YL(192) MMC_YL(256) MMC_YL(384) MMC_YL(512)
    MMC_YL(768) MMC_YL(1024) MMC_YL(2048)
#undef MMC_YL
    set_lds_one(e, (const void *)k_yfwd_tensor<T>, bytes);
    set_lds_one(e, (const void *)k_ymul_any<T, 0, false>, bytes);
    set_lds_one(e, (const void *)k_ymul_any<T, 1, false>, bytes);
    set_lds_one(e, (const void *)k_ymul_any<T, 2, false>, bytes);
    set_lds_one(e, (const void *)k_ymul_any<T, 3, false>, bytes);
    set_lds_one(e, (const void *)k_ymul_any<T, 0, true>, bytes);
    set_lds_one(e, (const void *)k_ymul_any<T, 1, true>, bytes);
    set_lds_one(e, (const void *)k_ymul_any<T, 2, true>, bytes);
    set_lds_one(e, (const void *)k_ymul_any<T, 3, true>, bytes);
#define MMC_YZL(NN) set_lds_one(e, (const void *)k_yzconv_n<T, NN, 2>, bytes); \
    set_lds_one(e, (const void *)k_yzconv_n<T, NN, 3>, bytes);                 \
    set_lds_one(e, (const void *)k_yzconv_n<T, NN, 4>, bytes);
    MMC_YZL(64) MMC_YZL(96) MMC_YZL(128) MMC_YZL(192) MMC_YZL(256) MMC_YZL(384) MMC_YZL(512)
    MMC_YZL(768) MMC_YZL(1024) MMC_YZL(2048)
#undef MMC_YZL
    return e;
}

#define MMC_INST(T)                                                                           \
    template hipError_t launch_zfwd<T>(hipStream_t, const Inst *, const Work *, int, int, int, \
                                       const cpx<T> *, cpx<T> *, const cpx<T> *, size_t, bool); \
    template hipError_t launch_zinv<T>(hipStream_t, const Inst *, const Work *, int, int,     \
                                       const cpx<T> *, cpx<T> *, const cpx<T> *, size_t, bool); \
    template hipError_t launch_ymul<T>(hipStream_t, const Inst *, const Work *, int, int, int,\
                                       cpx<T> *, const cpx<T> *, const cpx<T> *, size_t, bool, \
                                       bool);                                                  \
    template hipError_t launch_yzconv<T>(hipStream_t, const Inst *, const Work *, int, int, int,\
                                         cpx<T> *, const cpx<T> *, const cpx<T> *, size_t,     \
                                         bool);                                                \
    template hipError_t launch_yfwd_tensor<T>(hipStream_t, const Inst *, const Work *, int,   \
                                              const cpx<double> *, cpx<T> *,                  \
                                              const cpx<double> *, size_t);                   \
    template hipError_t set_lds_ypass<T>(int);
MMC_INST(float)
MMC_INST(double)
#undef MMC_INST

}  // namespace mmc
