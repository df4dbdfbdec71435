// mmc_device.h -- device helpers shared by the kernel translation units:
// Philox, the uniform map, LDS staging of global data, the field stencil and
// the Metropolis step, and the launch macro that records dispatch-packet
// events for timed launches.
#pragma once

#include "fft_fixed.h"
#include "fft_lds.h"
#include "mmc_internal.h"

#include <hip/hip_ext.h>

#include <cstdint>

namespace mmc {

// ============================================================================
// small helpers
// ============================================================================
// Neighbour of i at step d = +-1 along an axis of n cells: -1 past an open
// end or when the wrap lands on i itself (n = 1).
__device__ __forceinline__ int nbr(int i, int d, int n, int periodic)
{
    int j = i + d;
    if (j < 0 || j >= n) {
        if (!periodic) return -1;
        j = j < 0 ? j + n : j - n;   // j is -1 or n: (j + n) mod n without a division
    }
    return j == i ? -1 : j;
}

__device__ __forceinline__ int axis_color(int i, int n, int periodic)
{
    return (periodic && n >= 3 && (n & 1) && i == n - 1) ? 2 : (i & 1);
}

// Philox4x32-10 (Salmon et al. 2011).  Each round's two 32x32 -> 64-bit
// products are formed as one 64-bit product each (one v_mad_u64_u32 instead
// of separate high and low multiplies: 1.36x the throughput, same bits).
__device__ __forceinline__ uint4 philox10(uint4 c, uint32_t k0, uint32_t k1)
{
    #pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1,
                       (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// 23 random bits -> (2k+1)/2^24, exactly representable in FP32 and FP64.
template <typename T>
__device__ __forceinline__ T u01(uint32_t r)
{
    return T(2u * (r >> 9) + 1u) * T(1.0 / 16777216.0);
}

__device__ __forceinline__ void sincospi_t(float a, float *s, float *c) { sincospif(a, s, c); }
__device__ __forceinline__ void sincospi_t(double a, double *s, double *c) { sincospi(a, s, c); }
__device__ __forceinline__ float exp_t(float a) { return expf(a); }
__device__ __forceinline__ double exp_t(double a) { return exp(a); }
__device__ __forceinline__ float sqrt_t(float a) { return sqrtf(a); }
__device__ __forceinline__ double sqrt_t(double a) { return sqrt(a); }

// ============================================================================
// (b) FFT passes
// ============================================================================
extern __shared__ __align__(16) unsigned char g_smem[];

// Elements e = threadIdx.x + q*NT (q < ceil(MAXC/NT), e < count) of a
// global -> LDS copy: ld(e) reads element e into registers, st(e, v) writes
// it.  A rolled loop waits for each load before its LDS write, so a thread
// pays one global latency per element; here every load of a chunk is issued
// before the chunk's first write, so the latencies overlap.  Chunks hold at
// most 128 bytes of loaded values per thread.
template <int MAXC, int NT, typename V, typename LD, typename ST, typename MID>
__device__ __forceinline__ void staged_copy(int count, LD ld, ST st, MID mid);

template <int MAXC, int NT, typename V, typename LD, typename ST>
__device__ __forceinline__ void staged_copy(int count, LD ld, ST st)
{
    staged_copy<MAXC, NT, V>(count, ld, st, [] {});
}

// The same, with mid() run once after the first chunk's loads are issued
// and before their values are used: work that needs none of them (drawing
// random numbers, say) fills the loads' latency.
template <int MAXC, int NT, typename V, typename LD, typename ST, typename MID>
__device__ __forceinline__ void staged_copy(int count, LD ld, ST st, MID mid)
{
    constexpr int Q = (MAXC + NT - 1) / NT;
    constexpr int CHM = 128 / (int)sizeof(V) > 0 ? 128 / (int)sizeof(V) : 1;
    constexpr int CH = Q < CHM ? Q : CHM;
    #pragma unroll
    for (int q0 = 0; q0 < Q; q0 += CH) {
        V v[CH];
        #pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int e = threadIdx.x + (q0 + q) * NT;
            if (q0 + q < Q && e < count) v[q] = ld(e);
        }
        if (q0 == 0) mid();
        #pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int e = threadIdx.x + (q0 + q) * NT;
            if (q0 + q < Q && e < count) st(e, v[q]);
        }
    }
}

// The same for a count known only at run time: chunks of CH loads per
// thread (at most 128 bytes), one global latency per chunk.
template <typename V, int NT = kThreads, typename LD, typename ST>
__device__ __forceinline__ void staged_copy_dyn(int count, LD ld, ST st)
{
    constexpr int CH = 128 / (int)sizeof(V) > 0 ? 128 / (int)sizeof(V) : 1;
    for (int b = 0; b < count; b += CH * NT) {
        V v[CH];
        #pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int e = b + threadIdx.x + q * NT;
            if (e < count) v[q] = ld(e);
        }
        #pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int e = b + threadIdx.x + q * NT;
            if (e < count) st(e, v[q]);
        }
    }
}

// A staged copy split in two, for a few elements per thread (Q): the loads
// are issued by staged_load, the LDS writes done by staged_store, so loads
// of other data can be issued in between.  Vector-memory loads complete in
// issue order, so data needed first is loaded first.
template <int Q, typename V> struct Staged {
    V v[Q];
};

template <int MAXC, int NT, typename V, typename LD>
__device__ __forceinline__ Staged<(MAXC + NT - 1) / NT, V> staged_load(int count, LD ld)
{
    constexpr int Q = (MAXC + NT - 1) / NT;
    Staged<Q, V> r;
    #pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = threadIdx.x + q * NT;
        if (e < count) r.v[q] = ld(e);
    }
    return r;
}

template <int MAXC, int NT, typename V, typename ST>
__device__ __forceinline__ void staged_store(int count, const Staged<(MAXC + NT - 1) / NT, V> &r,
                                             ST st)
{
    constexpr int Q = (MAXC + NT - 1) / NT;
    #pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int e = threadIdx.x + q * NT;
        if (e < count) st(e, r.v[q]);
    }
}

// Twiddle table of a compile-time length N into LDS in two steps: the loads
// (load_twiddles) are issued before the kernel's data loads and the LDS
// writes (store_twiddles) done after them, so the table's latency overlaps
// the data's instead of preceding it.
template <typename T, int N, int NT = kThreads>
__device__ __forceinline__ Staged<(N + NT - 1) / NT, cpx<T>> load_twiddles(const cpx<T> *__restrict__ src)
{
    return staged_load<N, NT, cpx<T>>(N, [&](int e) { return src[e]; });
}

template <typename T, int N, int NT = kThreads>
__device__ __forceinline__ void store_twiddles(cpx<T> *dst, const Staged<(N + NT - 1) / NT, cpx<T>> &r)
{
    staged_store<N, NT, cpx<T>>(N, r, [&](int e, cpx<T> v) { dst[e] = v; });
}

// Twiddle table of a compile-time length N into LDS, loads batched.
template <typename T, int N, int NT = kThreads>
__device__ __forceinline__ void stage_twiddles(cpx<T> *dst, const cpx<T> *__restrict__ src)
{
    staged_copy<N, NT, cpx<T>>(
        N, [&](int e) { return src[e]; }, [&](int e, cpx<T> v) { dst[e] = v; });
}

// Two values loaded together by one thread of a staged copy.
template <typename V> struct Pair2 { V a, b; };

// Rows (y, y+1) of the y-contiguous spectrum at one (component, kx, z) as
// one 16-byte access.  Lanes of an x-pass tile read along kx, a line apart
// each, so halving the accesses halves the address work of those loads.
// Valid for FP32 when every pair is complete (even ny) and the instance's
// spectrum starts 16-byte aligned: pair16(in).
template <typename T>
__device__ __forceinline__ bool pair16(const Inst &in)
{
    return sizeof(T) == 4 && (in.ny & 1) == 0 && (in.a_off & 1) == 0;
}

template <typename T>
__device__ __forceinline__ Pair2<cpx<T>> load_pair16(const cpx<T> *p)
{
    const float4 f = *reinterpret_cast<const float4 *>(p);
    Pair2<cpx<T>> v;
    v.a = {T(f.x), T(f.y)};
    v.b = {T(f.z), T(f.w)};
    return v;
}

template <typename T>
__device__ __forceinline__ void store_pair16(cpx<T> *p, const cpx<T> &a, const cpx<T> &b)
{
    *reinterpret_cast<float4 *>(p) = make_float4(float(a.x), float(a.y), float(b.x), float(b.y));
}

// Copy a twiddle table into LDS (after the two ping-pong buffers); callers
// synchronise once data and twiddles are both loaded.
template <typename T>
__device__ __forceinline__ const cpx<T> *lds_twiddles(cpx<T> *dst, const cpx<T> *__restrict__ src,
                                                      int n)
{
    staged_copy_dyn<cpx<T>>(
        n, [&](int e) { return src[e]; }, [&](int e, cpx<T> v) { dst[e] = v; });
    return dst;
}
// ============================================================================
// (c)+(d) local field stencil and Metropolis update
// ============================================================================
template <typename T> struct V3 { T x, y, z; };

// Cell indices of one instance are 32-bit (the host keeps 3N below 2^32):
// cheaper than 64-bit arithmetic, whose multiply-adds read undefined high
// halves of registers that in-flight loads write, so the compiler waited for
// those loads.
// x + nx*(y + ny*z) with 24-bit products (full-rate multiply-adds; the host
// keeps nx and ny*nz below 2^24)
__device__ __forceinline__ uint32_t cell_index(uint32_t nx, uint32_t ny, int x, int y, int z)
{
    return (uint32_t)x + __umul24(nx, (uint32_t)y + __umul24(ny, (uint32_t)z));
}

template <typename T>
__device__ __forceinline__ V3<T> load_m(const T *m, uint32_t N, uint32_t i)
{
    return {m[i], m[N + i], m[2u * N + i]};
}

// Quotient a / b of small non-negative ints (a, b < 2^22) by a divisor known
// only at run time: a float reciprocal, rounded, gives the quotient within
// one, and one exact correction each way fixes it (about 8 instructions
// instead of an integer division's 20).
struct Div {
    int b;
    float rb;
    __device__ explicit Div(int d) : b(d), rb(1.0f / (float)d) {}
    __device__ __forceinline__ int operator()(int a) const
    {
        const int q = (int)((float)a * rb);
        const int r = a - __mul24(q, b);
        return q + (r >= b ? 1 : 0) - (r < 0 ? 1 : 0);
    }
};

// Physics constants of an instance in the working precision T: the FP64
// fields, or for FP32 their copies rounded once on the host (Inst::pf).
template <typename T> struct Ph {
    static __device__ __forceinline__ T lam(const Inst &in, int a) { return in.lam[a]; }
    static __device__ __forceinline__ T dmi(const Inst &in, int a) { return in.dmi[a]; }
    static __device__ __forceinline__ T kap(const Inst &in) { return in.kap; }
    static __device__ __forceinline__ T u(const Inst &in, int a) { return in.u[a]; }
    static __device__ __forceinline__ T b(const Inst &in, int a) { return in.b[a]; }
    static __device__ __forceinline__ T K0(const Inst &in, int k) { return in.K0[k]; }
};
template <> struct Ph<float> {
    static __device__ __forceinline__ float lam(const Inst &in, int a) { return in.pf.lam[a]; }
    static __device__ __forceinline__ float dmi(const Inst &in, int a) { return in.pf.dmi[a]; }
    static __device__ __forceinline__ float kap(const Inst &in) { return in.pf.kap; }
    static __device__ __forceinline__ float u(const Inst &in, int a) { return in.pf.u[a]; }
    static __device__ __forceinline__ float b(const Inst &in, int a) { return in.pf.b[a]; }
    static __device__ __forceinline__ float K0(const Inst &in, int k) { return in.pf.K0[k]; }
};

// Bond sums of one cell from its six neighbour values (a missing bond
// passes zeros), reduced units (field per Ms):
//   ex  = sum_a lam_a * sum_{j in nbrs_a} m_j      (exchange, neighbours only)
//   dmi = interfacial DMI bond field (x and y bonds)
template <typename T>
__device__ __forceinline__ void bond_sums(const Inst &in, const V3<T> &xp, const V3<T> &xm,
                                          const V3<T> &yp, const V3<T> &ym, const V3<T> &zp,
                                          const V3<T> &zm, V3<T> &ex, V3<T> &dmi)
{
    {   // x bonds
        const T lam = Ph<T>::lam(in, 0), dm = Ph<T>::dmi(in, 0);
        ex = {lam * (xp.x + xm.x), lam * (xp.y + xm.y), lam * (xp.z + xm.z)};
        dmi = {dm * (xp.z - xm.z), T(0), -(dm * (xp.x - xm.x))};
    }
    {   // y bonds
        const T lam = Ph<T>::lam(in, 1), dm = Ph<T>::dmi(in, 1);
        ex.x = fmad(lam, yp.x + ym.x, ex.x);
        ex.y = fmad(lam, yp.y + ym.y, ex.y);
        ex.z = fmad(lam, yp.z + ym.z, ex.z);
        dmi.y = dm * (yp.z - ym.z);
        dmi.z = fmad(-dm, yp.y - ym.y, dmi.z);
    }
    if (in.nz > 1) {   // z bonds: exchange only
        const T lam = Ph<T>::lam(in, 2);
        ex.x = fmad(lam, zp.x + zm.x, ex.x);
        ex.y = fmad(lam, zp.y + zm.y, ex.y);
        ex.z = fmad(lam, zp.z + zm.z, ex.z);
    }
}

// Neighbour sums of cell (x,y,z) read from global m, plus
//   cnt = sum_a lam_a * (#neighbours along a)       (Laplacian diagonal)
template <typename T>
__device__ __forceinline__ void stencil(const Inst &in, const T *m, int x, int y, int z,
                                        V3<T> &ex, V3<T> &dmi, T &cnt)
{
    const uint32_t N = (uint32_t)in.N;
    const int nx = in.nx, ny = in.ny, nz = in.nz;
    const V3<T> zero = {T(0), T(0), T(0)};
    V3<T> xp = zero, xm = zero, yp = zero, ym = zero, zp = zero, zm = zero;
    cnt = T(0);
    const uint32_t unx = nx, uny = ny;
    const uint32_t row = cell_index(unx, uny, 0, y, z);
    int j;
    if ((j = nbr(x, +1, nx, in.pbc[0])) >= 0) { xp = load_m(m, N, row + j); cnt += Ph<T>::lam(in, 0); }
    if ((j = nbr(x, -1, nx, in.pbc[0])) >= 0) { xm = load_m(m, N, row + j); cnt += Ph<T>::lam(in, 0); }
    if ((j = nbr(y, +1, ny, in.pbc[1])) >= 0) { yp = load_m(m, N, cell_index(unx, uny, x, j, z)); cnt += Ph<T>::lam(in, 1); }
    if ((j = nbr(y, -1, ny, in.pbc[1])) >= 0) { ym = load_m(m, N, cell_index(unx, uny, x, j, z)); cnt += Ph<T>::lam(in, 1); }
    if (nz > 1) {
        if ((j = nbr(z, +1, nz, in.pbc[2])) >= 0) { zp = load_m(m, N, cell_index(unx, uny, x, y, j)); cnt += Ph<T>::lam(in, 2); }
        if ((j = nbr(z, -1, nz, in.pbc[2])) >= 0) { zm = load_m(m, N, cell_index(unx, uny, x, y, j)); cnt += Ph<T>::lam(in, 2); }
    }
    bond_sums(in, xp, xm, yp, ym, zp, zm, ex, dmi);
}

// a.K0.a for the symmetric self tensor K0 = (xx, xy, xz, yy, yz, zz).
template <typename T>
__device__ __forceinline__ T quad_k0(const Inst &in, V3<T> a)
{
    const T k0 = Ph<T>::K0(in, 0), k1 = Ph<T>::K0(in, 1), k2 = Ph<T>::K0(in, 2);
    const T k3 = Ph<T>::K0(in, 3), k4 = Ph<T>::K0(in, 4), k5 = Ph<T>::K0(in, 5);
    const T kx = fmad(k0, a.x, fmad(T(2) * k1, a.y, T(2) * k2 * a.z));
    const T ky = fmad(k3, a.y, T(2) * k4 * a.z);
    return fmad(a.x, kx, fmad(a.y, ky, k5 * a.z * a.z));
}

template <typename T> struct Move {
    V3<T> mn;   // trial direction
    T u;        // acceptance draw
};

// Trial direction (uniform on the sphere) and acceptance draw of a cell from
// Philox counter (cell, sweep, stage, 0): a function of the counter alone,
// so it can be drawn before the cell's inputs arrive.
template <typename T>
__device__ __forceinline__ Move<T> trial_move(const Inst &in, uint32_t cell, uint32_t sweep,
                                              uint32_t stage)
{
    const uint4 rr = philox10(make_uint4(cell, sweep, stage, 0u), in.key[0], in.key[1]);
    const T zt = T(1) - T(2) * u01<T>(rr.x);
    const T st = sqrt_t(max(T(0), T(1) - zt * zt));
    T sn, cs;
    sincospi_t(T(2) * u01<T>(rr.y), &sn, &cs);
    return {{st * cs, st * sn, zt}, u01<T>(rr.z)};
}

// Energy change of moving cell m_i to mn (see trial_de).
template <typename T>
__device__ __forceinline__ T move_de(const Inst &in, const V3<T> &mi, const V3<T> &ex,
                                     const V3<T> &dmi, const V3<T> &hi, int use_demag,
                                     const V3<T> &mn)
{
    // with d = m' - m and the frozen demag field h (which includes the
    // cell's own contribution K0 m):
    //   -d.(h - K0 m) - [m'K0m' - mK0m]/2 = -d.h - d.K0.d/2
    // (K0 symmetric), one quadratic form instead of three
    V3<T> g = {ex.x + dmi.x + Ph<T>::b(in, 0), ex.y + dmi.y + Ph<T>::b(in, 1),
               ex.z + dmi.z + Ph<T>::b(in, 2)};
    if (use_demag) {
        g.x += hi.x;
        g.y += hi.y;
        g.z += hi.z;
    }
    const V3<T> d = {mn.x - mi.x, mn.y - mi.y, mn.z - mi.z};
    T dE = -fmad(d.x, g.x, fmad(d.y, g.y, d.z * g.z));
    const T ux = Ph<T>::u(in, 0), uy = Ph<T>::u(in, 1), uz = Ph<T>::u(in, 2);
    const T pu = fmad(mn.x, ux, fmad(mn.y, uy, mn.z * uz));
    const T ou = fmad(mi.x, ux, fmad(mi.y, uy, mi.z * uz));
    dE = fmad(-Ph<T>::kap(in), fmad(pu, pu, -(ou * ou)), dE);
    if (use_demag) dE = fmad(T(-0.5), quad_k0(in, d), dE);
    return dE;
}

// Trial move of one cell (counter (cell, sweep, stage, 0)) and its energy
// change, in units of mu0 Ms^2 V:
//   dE = -d.(ex + dmi + b + h - K0 m) - kap[(m'.u)^2 - (m.u)^2] - [m'K0m' - mK0m]/2
// with d = m' - m.
template <typename T>
__device__ __forceinline__ T trial_de(const Inst &in, const V3<T> &mi, const V3<T> &ex,
                                      const V3<T> &dmi, const V3<T> &hi, int use_demag,
                                      uint32_t cell, uint32_t sweep, uint32_t stage, V3<T> &mn)
{
    mn = trial_move<T>(in, cell, sweep, stage).mn;
    return move_de(in, mi, ex, dmi, hi, use_demag, mn);
}

// Metropolis decision for a drawn move: accept if dE <= 0 or
// u < exp(-dE * beta), beta = mu0 Ms^2 V / (kB T).
template <typename T>
__device__ __forceinline__ bool metro_accept(const Inst &in, const V3<T> &mi, const V3<T> &ex,
                                             const V3<T> &dmi, const V3<T> &hi, int use_demag,
                                             const Move<T> &mv, T bt)
{
    const T dE = move_de(in, mi, ex, dmi, hi, use_demag, mv.mn);
    return dE <= T(0) || mv.u < exp_t(-dE * bt);
}

// Metropolis decision for one cell, move drawn here; mn = trial direction.
template <typename T>
__device__ __forceinline__ bool metro_cell(const Inst &in, const V3<T> &mi, const V3<T> &ex,
                                           const V3<T> &dmi, const V3<T> &hi, int use_demag,
                                           uint32_t cell, uint32_t sweep, uint32_t stage, T bt,
                                           V3<T> &mn)
{
    const Move<T> mv = trial_move<T>(in, cell, sweep, stage);
    mn = mv.mn;
    return metro_accept(in, mi, ex, dmi, hi, use_demag, mv, bt);
}

// The six neighbour values of cell (x,y,z) as stencil() reads them (zero for
// a missing bond), loaded but not yet summed.
template <typename T> struct Nbrs {
    V3<T> xp, xm, yp, ym, zp, zm;
};

template <typename T>
__device__ __forceinline__ Nbrs<T> gather_nbrs(const Inst &in, const T *m, int x, int y, int z)
{
    const uint32_t N = (uint32_t)in.N;
    const int nx = in.nx, ny = in.ny, nz = in.nz;
    const V3<T> zero = {T(0), T(0), T(0)};
    Nbrs<T> b = {zero, zero, zero, zero, zero, zero};
    const uint32_t unx = nx, uny = ny;
    const uint32_t row = cell_index(unx, uny, 0, y, z);
    int j;
    if ((j = nbr(x, +1, nx, in.pbc[0])) >= 0) b.xp = load_m(m, N, row + j);
    if ((j = nbr(x, -1, nx, in.pbc[0])) >= 0) b.xm = load_m(m, N, row + j);
    if ((j = nbr(y, +1, ny, in.pbc[1])) >= 0) b.yp = load_m(m, N, cell_index(unx, uny, x, j, z));
    if ((j = nbr(y, -1, ny, in.pbc[1])) >= 0) b.ym = load_m(m, N, cell_index(unx, uny, x, j, z));
    if (nz > 1) {
        if ((j = nbr(z, +1, nz, in.pbc[2])) >= 0) b.zp = load_m(m, N, cell_index(unx, uny, x, y, j));
        if ((j = nbr(z, -1, nz, in.pbc[2])) >= 0) b.zm = load_m(m, N, cell_index(unx, uny, x, y, j));
    }
    return b;
}

// Metropolis update of checkerboard cell i from its gathered neighbours (of
// the other colour, so no workgroup writes them during the pass) and a drawn
// move: writes global m when accepted, returns the cell's new value.
template <typename T>
__device__ __forceinline__ V3<T> metro_apply(const Inst &in, T *m, long i, const V3<T> &mi,
                                             const Nbrs<T> &nb, const V3<T> &hi,
                                             const Move<T> &mv, T bt)
{
    V3<T> ex, dmi;
    bond_sums(in, nb.xp, nb.xm, nb.yp, nb.ym, nb.zp, nb.zm, ex, dmi);
    if (metro_accept(in, mi, ex, dmi, hi, 1, mv, bt)) {
        const long N = in.N;
        m[i] = mv.mn.x;
        m[N + i] = mv.mn.y;
        m[2 * N + i] = mv.mn.z;
        return mv.mn;
    }
    return mi;
}

// Cells per thread whose inputs k_metro_xfwd fetches before its LDS phase:
// all QC when they fit (two FP32 / one FP64 cell), else none.
template <typename T, int QC>
__host__ __device__ constexpr int prefetch_cells()
{
    return QC <= (sizeof(T) == 4 ? 2 : 1) ? QC : 0;
}
// Cells whose inputs are loaded together when not prefetched.
template <typename T>
__host__ __device__ constexpr int gather_chunk()
{
    return sizeof(T) == 4 ? 2 : 1;
}
// Direct DFT of length P <= kZDirect held in registers: nin inputs (zero
// above), nout outputs, forward or inverse (conjugate twiddles).
template <typename T, bool INV>
__device__ __forceinline__ void zdft_regs(const cpx<T> (&v)[kZDirect], cpx<T> (&out)[kZDirect],
                                          int nin, int nout, int P,
                                          const cpx<T> *__restrict__ tw)
{
    #pragma unroll
    for (int k = 0; k < kZDirect; ++k) {
        cpx<T> acc = v[0];
        #pragma unroll
        for (int z = 1; z < kZDirect; ++z) {
            if (z < nin && k < nout) {
                cpx<T> w = tw[(k * z) % P];
                if (INV) w.y = -w.y;
                acc = cadd(acc, cmul(v[z], w));
            }
        }
        out[k] = acc;
    }
}

// Direct DFT along z of one y column (element z at p[z*stride]): nin inputs
// (zero above), nout outputs, length P <= kZDirect, forward or inverse
// (conjugate twiddles).  In place safe: all inputs are read first.
template <typename T, bool INV>
__device__ __forceinline__ void zdft_column(const cpx<T> *src, cpx<T> *dst, int stride,
                                            int nin, int nout, int P,
                                            const cpx<T> *__restrict__ tw)
{
    cpx<T> v[kZDirect], out[kZDirect];
    #pragma unroll
    for (int z = 0; z < kZDirect; ++z)
        v[z] = z < nin ? src[(long)z * stride] : cpx<T>{T(0), T(0)};
    zdft_regs<T, INV>(v, out, nin, nout, P, tw);
    #pragma unroll
    for (int k = 0; k < kZDirect; ++k)
        if (k < nout) dst[(long)k * stride] = out[k];
}

// ============================================================================
// launch helpers
// ============================================================================
// Events for the next launch of this thread (set_launch_events); cleared by
// post() after every launch wrapper.
extern thread_local hipEvent_t g_ev_a, g_ev_b;

inline hipError_t post()
{
    g_ev_a = g_ev_b = nullptr;
    return hipGetLastError();
}

// Timed launches record their events in the dispatch packet
// (hipExtLaunchKernelGGL); untimed ones are plain launches, which every
// profiler mode intercepts.
#define MMC_LAUNCH(kern, grid, block, lds, st, ...)                                              \
    do {                                                                                         \
        if (g_ev_a)                                                                              \
            hipExtLaunchKernelGGL(kern, grid, block, (std::uint32_t)(lds), st, g_ev_a, g_ev_b, 0u, \
                                  __VA_ARGS__);                                                  \
        else                                                                                     \
            hipLaunchKernelGGL(kern, grid, block, (std::uint32_t)(lds), st, __VA_ARGS__);        \
    } while (0)

// Fixed x-pass variants: sizes x {1 row pair, kXRows<T> row pairs}.
#define MMC_XFIXED(M, T, R) M(T, 128, R) M(T, 192, R) M(T, 256, R) M(T, 384, R) \
    M(T, 512, R) M(T, 1024, R) M(T, 2048, R)
// fused x + Metropolis variants: every fixed x length, one row pair or the
// full tile fused_rows(F, N)
#define MMC_FUSED(M, T, F) M(T, 128, 1) M(T, 192, 1) M(T, 256, 1) M(T, 384, 1) M(T, 512, 1)   \
    M(T, 1024, 1) M(T, 2048, 1) M(T, 128, fused_rows(F, 128)) M(T, 192, fused_rows(F, 192))    \
    M(T, 256, fused_rows(F, 256)) M(T, 384, fused_rows(F, 384)) M(T, 512, fused_rows(F, 512))  \
    M(T, 1024, fused_rows(F, 1024))

// Each kernel translation unit raises the dynamic-LDS limit of its own FFT
// kernels (set_lds_limit calls them all).
template <typename T> hipError_t set_lds_xpass(int bytes);
template <typename T> hipError_t set_lds_ypass(int bytes);
template <typename T> hipError_t set_lds_fused_inv(int bytes);
template <typename T> hipError_t set_lds_fused_fwd(int bytes);
template <typename T> hipError_t set_lds_rows(int bytes);
template <typename T> hipError_t set_lds_metro(int bytes);

// hipFuncSetAttribute of one kernel, first failure kept in e.
inline void set_lds_one(hipError_t &e, const void *f, int bytes)
{
    hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (r != hipSuccess) e = r;
}

}  // namespace mmc
