// k_metro.hip -- Metropolis update per colour class and the tiled
// two-colour variant (mcKernel2, src/mc_sampling.h:27-41); energy terms and
// <m> with compensated FP64 reductions (calEnergyKernel + thrust::reduce,
// src/main.cu:902-956); effective field (effectiveHKernel,
// src/effective_H.h:25-33); energy density; trial dE; the Philox start
// state; the launch-event hooks and the LDS limit of every kernel.
#include "mmc_device.h"

namespace mmc {

// One colour class, one thread per cell of the class (mcKernel2,
// src/mc_sampling.h:27-41); used when a periodic axis has odd length
// (three classes) and by the tiled kernel's fallback.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_metro(const Inst *__restrict__ I,
                                                    const Work *__restrict__ W,
                                                    T *__restrict__ mb,
                                                    const T *__restrict__ hb,
                                                    const double *__restrict__ beta,
                                                    const uint32_t *__restrict__ ctl,
                                                    uint32_t sweep_off, int color,
                                                    int use_demag)
{
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const int nx = in.nx, ny = in.ny;
    int x, y, z;
    long i;
    if (in.ncolors == 2) {
        // slot -> the colour's cell in its row: x = 2j + ((y + z + color) & 1)
        const int h2 = (nx + 1) >> 1;
        const long sl = (long)w.a + threadIdx.x;
        const long r = sl / h2;
        if (r >= (long)ny * in.nz) return;
        y = (int)(r % ny);
        z = (int)(r / ny);
        x = 2 * (int)(sl - r * h2) + ((y + z + color) & 1);
        if (x >= nx) return;
        i = x + (long)nx * r;
    } else {
        i = (long)w.a + threadIdx.x;
        if (i >= in.N) return;
        x = (int)(i % nx);
        const long r = i / nx;
        y = (int)(r % ny);
        z = (int)(r / ny);
        const int col = (axis_color(x, nx, in.pbc[0]) + axis_color(y, ny, in.pbc[1]) +
                         axis_color(z, in.nz, in.pbc[2])) % in.ncolors;
        if (col != color) return;
    }
    T *m = mb + in.m_off;
    const long N = in.N;
    const V3<T> mi = load_m(m, N, i);
    V3<T> ex, dmi;
    T cnt;
    stencil(in, m, x, y, z, ex, dmi, cnt);
    V3<T> hi = {T(0), T(0), T(0)};
    if (use_demag) hi = load_m(hb + in.h_off, N, i);
    V3<T> mn;
    if (metro_cell(in, mi, ex, dmi, hi, use_demag, (uint32_t)i, ctl[1] + sweep_off, ctl[0],
                   T(beta[w.inst]), mn)) {
        m[i] = mn.x;
        m[N + i] = mn.y;
        m[2 * N + i] = mn.z;
    }
}

// Both colour classes of a checkerboard in one pass.  Item (inst, x0, y0):
// an x-y tile of mtx x mty cells through all z layers.  LDS holds m with a
// two-cell x-y halo and h with a one-cell halo (periodic axes wrap, open
// edges are absent).  Colour 0 is updated on the tile plus a one-cell ring
// (the ring cells are the neighbours colour 1 needs; their updates repeat
// exactly what the neighbouring tile computes, same inputs and Philox
// counter), then colour 1 on the tile, and the tile is written back once.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_metro_tile(const Inst *__restrict__ I,
                                                         const Work *__restrict__ W,
                                                         T *__restrict__ mb,
                                                         const T *__restrict__ hb,
                                                         const double *__restrict__ beta,
                                                         const uint32_t *__restrict__ ctl,
                                                         uint32_t sweep_off, int use_demag)
{
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const int nx = in.nx, ny = in.ny, nz = in.nz;
    const int x0 = w.a, y0 = w.b;
    const int TX = in.mtx, TY = in.mty;
    const int MW = TX + 4, MH = TY + 4;          // m region: halo 2
    const int HW = TX + 2, HH = TY + 2;          // h region: halo 1
    const int mplane = MW * MH * nz, hplane = HW * HH * nz;
    T *sm = reinterpret_cast<T *>(g_smem);       // [3][nz][MH][MW]
    T *sh = sm + 3 * mplane;                     // [3][nz][HH][HW]
    const long N = in.N;
    const T *m = mb + in.m_off;
    const T *h = hb + in.h_off;
    const int px_ = in.pbc[0], py_ = in.pbc[1];
    // wrap a raw coordinate into [0,n) (periodic) or flag it absent (-1)
    auto wrap = [](int g, int n, int periodic) -> int {
        if (g >= 0 && g < n) return g;
        if (!periodic) return -1;
        if (g < 0 && g + n >= 0) return g + n;
        if (g >= n && g - n < n) return g - n;
        g %= n;
        return g < 0 ? g + n : g;
    };
    for (int e = threadIdx.x; e < mplane; e += kThreads) {
        const int xx = e % MW, t = e / MW, yy = t % MH, z = t / MH;
        const int gx = wrap(x0 - 2 + xx, nx, px_);
        const int gy = wrap(y0 - 2 + yy, ny, py_);
        V3<T> v = {T(0), T(0), T(0)};
        if (gx >= 0 && gy >= 0) v = load_m(m, N, gx + (long)nx * (gy + (long)ny * z));
        sm[e] = v.x;
        sm[mplane + e] = v.y;
        sm[2 * mplane + e] = v.z;
    }
    if (use_demag)
        for (int e = threadIdx.x; e < hplane; e += kThreads) {
            const int xx = e % HW, t = e / HW, yy = t % HH, z = t / HH;
            const int gx = wrap(x0 - 1 + xx, nx, px_);
            const int gy = wrap(y0 - 1 + yy, ny, py_);
            V3<T> v = {T(0), T(0), T(0)};
            if (gx >= 0 && gy >= 0) v = load_m(h, N, gx + (long)nx * (gy + (long)ny * z));
            sh[e] = v.x;
            sh[hplane + e] = v.y;
            sh[2 * hplane + e] = v.z;
        }
    __syncthreads();
    const uint32_t sweep = ctl[1] + sweep_off, stage = ctl[0];
    const T bt = T(beta[w.inst]);
    // a bond along an axis exists unless the cell sits on an open edge; a
    // periodic axis of length 1 has none (the cell would be its own neighbour)
    const bool xper = px_ && nx > 1, yper = py_ && ny > 1;
    // update the cell at m-region offset o (position px,py,z); global (gx,gy,z)
    auto update = [&](int o, int gx, int gy, int z) {
        const V3<T> zero = {T(0), T(0), T(0)};
        V3<T> xp = zero, xm = zero, yp = zero, ym = zero, zp = zero, zm = zero;
        auto at = [&](int q) -> V3<T> { return {sm[q], sm[mplane + q], sm[2 * mplane + q]}; };
        if (xper || gx + 1 < nx) xp = at(o + 1);
        if (xper || gx > 0) xm = at(o - 1);
        if (yper || gy + 1 < ny) yp = at(o + MW);
        if (yper || gy > 0) ym = at(o - MW);
        if (nz > 1) {
            int j;
            if ((j = nbr(z, +1, nz, in.pbc[2])) >= 0) zp = at(o + (j - z) * MW * MH);
            if ((j = nbr(z, -1, nz, in.pbc[2])) >= 0) zm = at(o + (j - z) * MW * MH);
        }
        V3<T> ex, dmi;
        bond_sums(in, xp, xm, yp, ym, zp, zm, ex, dmi);
        const V3<T> mi = at(o);
        V3<T> hi = zero;
        if (use_demag) {
            // h region is the m region shrunk by one cell on every side
            const int ph = o % MW, rest = o / MW, qy = rest % MH, qz = rest / MH;
            const int oh = (ph - 1) + HW * ((qy - 1) + HH * qz);
            hi = {sh[oh], sh[hplane + oh], sh[2 * hplane + oh]};
        }
        V3<T> mn;
        const uint32_t cell = (uint32_t)(gx + (long)nx * (gy + (long)ny * z));
        if (metro_cell(in, mi, ex, dmi, hi, use_demag, cell, sweep, stage, bt, mn)) {
            sm[o] = mn.x;
            sm[mplane + o] = mn.y;
            sm[2 * mplane + o] = mn.z;
        }
    };
    // colour 0 on the tile and its one-cell ring: in ring row (yy, z) the
    // colour-0 positions are xx = 2k + ((x0 + y0 + yy + z) & 1), xx < HW
    {
        const int hc = (HW + 1) >> 1;
        for (int e = threadIdx.x; e < hc * HH * nz; e += kThreads) {
            const int k = e % hc, r = e / hc, yy = r % HH, z = r / HH;
            const int xx = 2 * k + ((x0 + y0 + yy + z) & 1);
            if (xx >= HW) continue;
            const int gx = wrap(x0 - 1 + xx, nx, px_);
            const int gy = wrap(y0 - 1 + yy, ny, py_);
            if (gx < 0 || gy < 0) continue;
            update((xx + 1) + MW * ((yy + 1) + MH * z), gx, gy, z);
        }
    }
    __syncthreads();
    // colour 1 on the tile: xx = 2k + ((x0 + y0 + yy + z + 1) & 1), xx < tw
    const int tw = min(TX, nx - x0), th = min(TY, ny - y0);
    {
        const int hc = (tw + 1) >> 1;
        for (int e = threadIdx.x; e < hc * th * nz; e += kThreads) {
            const int k = e % hc, r = e / hc, yy = r % th, z = r / th;
            const int xx = 2 * k + ((x0 + y0 + yy + z + 1) & 1);
            if (xx >= tw) continue;
            update((xx + 2) + MW * ((yy + 2) + MH * z), x0 + xx, y0 + yy, z);
        }
    }
    __syncthreads();
    T *mo = mb + in.m_off;
    for (int e = threadIdx.x; e < tw * th * nz; e += kThreads) {
        const int xx = e % tw, t = e / tw, yy = t % th, z = t / th;
        const long gi = (x0 + xx) + (long)nx * ((y0 + yy) + (long)ny * z);
        const int o = (xx + 2) + MW * ((yy + 2) + MH * z);
        mo[gi] = sm[o];
        mo[N + gi] = sm[mplane + o];
        mo[2 * N + gi] = sm[2 * mplane + o];
    }
}

// ============================================================================
// (e) energy terms and <m>, compensated FP64 reductions
// ============================================================================
struct DD {
    double hi, lo;
};

__device__ __forceinline__ DD dd_add(DD a, DD b)
{
    const double s = a.hi + b.hi;
    const double v = s - a.hi;
    double e = (a.hi - (s - v)) + (b.hi - v);
    e += a.lo + b.lo;
    const double hi = s + e;
    return {hi, e - (hi - s)};
}

__device__ __forceinline__ DD dd_shfl_down(DD a, int off)
{
    return {__shfl_down(a.hi, off, 64), __shfl_down(a.lo, off, 64)};
}

// Fixed-order block reduction of NQ double-double quantities; result valid
// in thread 0.
template <int NQ>
__device__ __forceinline__ void block_reduce(DD *v)
{
    __shared__ DD red[kThreads / 64][NQ];
    #pragma unroll
    for (int q = 0; q < NQ; ++q)
        for (int off = 32; off > 0; off >>= 1) v[q] = dd_add(v[q], dd_shfl_down(v[q], off));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int q = 0; q < NQ; ++q) red[wid][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int q = 0; q < NQ; ++q) {
            DD a = red[0][q];
            for (int k = 1; k < kThreads / 64; ++k) a = dd_add(a, red[k][q]);
            v[q] = a;
        }
}

// Per-cell energy densities (units mu0 Ms^2):
//   e_ex = -m.(ex - cnt m)/2, e_dmi = -m.dmi/2, e_anis = -kap (m.u)^2,
//   e_zee = -m.b, e_demag = -m.h/2;  plus m itself.
// Energy terms of cell i (units mu0 Ms^2 V): e[0] exchange, e[1] DMI,
// e[2] anisotropy, e[3] Zeeman, e[4] demag (h of the current state); mi
// returns the cell's state.
template <typename T>
__device__ __forceinline__ void cell_energy(const Inst &in, const T *mb, const T *hb, long i,
                                            int use_demag, double e[5], V3<T> &mi)
{
    const int nx = in.nx, ny = in.ny;
    const int x = (int)(i % nx);
    const long r = i / nx;
    const int y = (int)(r % ny);
    const int z = (int)(r / ny);
    const T *m = mb + in.m_off;
    const long N = in.N;
    mi = load_m(m, N, i);
    V3<T> ex, dmi;
    T cnt;
    stencil(in, m, x, y, z, ex, dmi, cnt);
    const T hex_x = ex.x - cnt * mi.x, hex_y = ex.y - cnt * mi.y, hex_z = ex.z - cnt * mi.z;
    const T mu = mi.x * Ph<T>::u(in, 0) + mi.y * Ph<T>::u(in, 1) + mi.z * Ph<T>::u(in, 2);
    e[0] = double(T(-0.5) * (mi.x * hex_x + mi.y * hex_y + mi.z * hex_z));
    e[1] = double(T(-0.5) * (mi.x * dmi.x + mi.y * dmi.y + mi.z * dmi.z));
    e[2] = double(-Ph<T>::kap(in) * mu * mu);
    e[3] = double(-(mi.x * Ph<T>::b(in, 0) + mi.y * Ph<T>::b(in, 1) + mi.z * Ph<T>::b(in, 2)));
    e[4] = 0.0;
    if (use_demag) {
        const T *h = hb + in.h_off;
        e[4] = double(T(-0.5) * (mi.x * h[i] + mi.y * h[N + i] + mi.z * h[2 * N + i]));
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_energy(const Inst *__restrict__ I,
                                                     const Work *__restrict__ W,
                                                     const T *__restrict__ mb,
                                                     const T *__restrict__ hb,
                                                     double *__restrict__ part,
                                                     int use_demag)
{
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const long i = (long)w.a + threadIdx.x;
    DD v[8];
    #pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = {0.0, 0.0};
    if (i < in.N) {
        double e[5];
        V3<T> mi;
        cell_energy(in, mb, hb, i, use_demag, e, mi);
        #pragma unroll
        for (int q = 0; q < 5; ++q) v[q].hi = e[q];
        v[5].hi = double(mi.x);
        v[6].hi = double(mi.y);
        v[7].hi = double(mi.z);
    }
    block_reduce<8>(v);
    if (threadIdx.x == 0)
        for (int q = 0; q < 8; ++q) {
            part[(long)blockIdx.x * 16 + 2 * q] = v[q].hi;
            part[(long)blockIdx.x * 16 + 2 * q + 1] = v[q].lo;
        }
}

// One workgroup per instance; items first[inst] .. first[inst+1]-1.
__global__ __launch_bounds__(kThreads) void k_energy_final(const int *__restrict__ first,
                                                           const double *__restrict__ part,
                                                           double *__restrict__ out)
{
    const int inst = blockIdx.x;
    const int b0 = first[inst], b1 = first[inst + 1];
    DD v[8];
    #pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = {0.0, 0.0};
    for (int b = b0 + threadIdx.x; b < b1; b += kThreads)
        #pragma unroll
        for (int q = 0; q < 8; ++q)
            v[q] = dd_add(v[q], DD{part[(long)b * 16 + 2 * q], part[(long)b * 16 + 2 * q + 1]});
    block_reduce<8>(v);
    if (threadIdx.x == 0)
        for (int q = 0; q < 8; ++q) out[inst * 8 + q] = v[q].hi + v[q].lo;
}

// Effective field in A/m of the selected terms (bits: 1 exch, 2 dmi,
// 4 anis, 8 zeeman, 16 demag).  Output [3][N] doubles of the item's instance.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_field(const Inst *__restrict__ I,
                                                    const Work *__restrict__ W,
                                                    const T *__restrict__ mb,
                                                    const T *__restrict__ hb, int terms,
                                                    double *__restrict__ H)
{
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const long i = (long)w.a + threadIdx.x;
    if (i >= in.N) return;
    const int nx = in.nx, ny = in.ny;
    const int x = (int)(i % nx);
    const long r = i / nx;
    const int y = (int)(r % ny);
    const int z = (int)(r / ny);
    const T *m = mb + in.m_off;
    const long N = in.N;
    const V3<T> mi = load_m(m, N, i);
    V3<T> ex, dmi;
    T cnt;
    stencil(in, m, x, y, z, ex, dmi, cnt);
    T hx = T(0), hy = T(0), hz = T(0);
    if (terms & 1) { hx += ex.x - cnt * mi.x; hy += ex.y - cnt * mi.y; hz += ex.z - cnt * mi.z; }
    if (terms & 2) { hx += dmi.x; hy += dmi.y; hz += dmi.z; }
    if (terms & 4) {
        const T mu = mi.x * Ph<T>::u(in, 0) + mi.y * Ph<T>::u(in, 1) + mi.z * Ph<T>::u(in, 2);
        const T f = T(2) * Ph<T>::kap(in) * mu;
        hx += f * Ph<T>::u(in, 0); hy += f * Ph<T>::u(in, 1); hz += f * Ph<T>::u(in, 2);
    }
    if (terms & 8) { hx += Ph<T>::b(in, 0); hy += Ph<T>::b(in, 1); hz += Ph<T>::b(in, 2); }
    if ((terms & 16) && in.cal_demag) {
        const T *h = hb + in.h_off;
        hx += h[i]; hy += h[N + i]; hz += h[2 * N + i];
    }
    H[i] = in.ms * double(hx);
    H[N + i] = in.ms * double(hy);
    H[2 * N + i] = in.ms * double(hz);
}

// Energy density [J/m^3] of every cell, the selected terms (bits 1 exch,
// 2 dmi, 4 anis, 8 zeeman, 16 demag) summed in that order.  Output N doubles.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_energy_density(const Inst *__restrict__ I,
                                                             const Work *__restrict__ W,
                                                             const T *__restrict__ mb,
                                                             const T *__restrict__ hb, int terms,
                                                             double *__restrict__ out)
{
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const long i = (long)w.a + threadIdx.x;
    if (i >= in.N) return;
    double e[5];
    V3<T> mi;
    cell_energy(in, mb, hb, i, in.cal_demag, e, mi);
    double t = 0.0;
    #pragma unroll
    for (int q = 0; q < 5; ++q)
        if (terms & (1 << q)) t += e[q];
    out[i] = t * (in.e_unit / (in.d[0] * in.d[1] * in.d[2]));
}

// Energy change [J] of every cell's trial move (Philox counter (cell, sweep,
// stage, 0)) against the current state and its demag field h, the move not
// applied: the dE that the Metropolis kernels test.  Output N doubles.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_delta_e(const Inst *__restrict__ I,
                                                      const Work *__restrict__ W,
                                                      const T *__restrict__ mb,
                                                      const T *__restrict__ hb, uint32_t stage,
                                                      uint32_t sweep, double *__restrict__ out)
{
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const long i = (long)w.a + threadIdx.x;
    if (i >= in.N) return;
    const int nx = in.nx, ny = in.ny;
    const int x = (int)(i % nx);
    const long r = i / nx;
    const int y = (int)(r % ny);
    const int z = (int)(r / ny);
    const T *m = mb + in.m_off;
    const long N = in.N;
    const V3<T> mi = load_m(m, N, i);
    V3<T> ex, dmi, mn;
    T cnt;
    stencil(in, m, x, y, z, ex, dmi, cnt);
    V3<T> hi = {T(0), T(0), T(0)};
    if (in.cal_demag) hi = load_m(hb + in.h_off, N, i);
    const T dE = trial_de(in, mi, ex, dmi, hi, in.cal_demag, (uint32_t)i, sweep, stage, mn);
    out[i] = in.e_unit * double(dE);
}

// Uniform-on-sphere start state, Philox counter (cell, 0, 0, 1).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_init(const Inst *__restrict__ I,
                                                   const Work *__restrict__ W,
                                                   T *__restrict__ mb)
{
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const long i = (long)w.a + threadIdx.x;
    if (i >= in.N) return;
    const uint4 rr = philox10(make_uint4((uint32_t)i, 0u, 0u, 1u), in.key[0], in.key[1]);
    const T zt = T(1) - T(2) * u01<T>(rr.x);
    const T st = sqrt_t(max(T(0), T(1) - zt * zt));
    T sn, cs;
    sincospi_t(T(2) * u01<T>(rr.y), &sn, &cs);
    T *m = mb + in.m_off;
    m[i] = st * cs;
    m[in.N + i] = st * sn;
    m[2 * in.N + i] = zt;
}

__global__ void k_set_ctl(uint32_t *ctl, uint32_t stage, uint32_t sweep_base)
{
    if (threadIdx.x == 0) {
        ctl[0] = stage;
        ctl[1] = sweep_base;
    }
}

thread_local hipEvent_t g_ev_a = nullptr, g_ev_b = nullptr;

void set_launch_events(hipEvent_t a, hipEvent_t b)
{
    g_ev_a = a;
    g_ev_b = b;
}

template <typename T>
hipError_t launch_metro(hipStream_t st, const Inst *I, const Work *W, int n, T *m, const T *h,
                        const double *beta, const uint32_t *ctl, uint32_t sweep_off, int color,
                        int use_demag)
{
    if (n <= 0) return hipSuccess;
    MMC_LAUNCH(k_metro<T>, dim3(n), dim3(kThreads), 0, st, I, W, m, h, beta, ctl,
                       sweep_off, color, use_demag);
    return post();
}

template <typename T>
hipError_t launch_metro_tile(hipStream_t st, const Inst *I, const Work *W, int n, T *m, const T *h,
                             const double *beta, const uint32_t *ctl, uint32_t sweep_off,
                             int use_demag, size_t lds)
{
    if (n <= 0) return hipSuccess;
    MMC_LAUNCH(k_metro_tile<T>, dim3(n), dim3(kThreads), lds, st, I, W, m, h, beta, ctl, sweep_off,
               use_demag);
    return post();
}

template <typename T>
hipError_t launch_energy(hipStream_t st, const Inst *I, const Work *W, int n, const T *m,
                         const T *h, double *partials, int use_demag)
{
    if (n <= 0) return hipSuccess;
    MMC_LAUNCH(k_energy<T>, dim3(n), dim3(kThreads), 0, st, I, W, m, h, partials,
                       use_demag);
    return post();
}

hipError_t launch_energy_final(hipStream_t st, const int *first, int ninst,
                               const double *partials, double *out)
{
    if (ninst <= 0) return hipSuccess;
    MMC_LAUNCH(k_energy_final, dim3(ninst), dim3(kThreads), 0, st, first, partials, out);
    return post();
}

template <typename T>
hipError_t launch_field(hipStream_t st, const Inst *I, const Work *W, int n, const T *m,
                        const T *h, int terms, double *H)
{
    if (n <= 0) return hipSuccess;
    MMC_LAUNCH(k_field<T>, dim3(n), dim3(kThreads), 0, st, I, W, m, h, terms, H);
    return post();
}

template <typename T>
hipError_t launch_energy_density(hipStream_t st, const Inst *I, const Work *W, int n, const T *m,
                                 const T *h, int terms, double *out)
{
    if (n <= 0) return hipSuccess;
    MMC_LAUNCH(k_energy_density<T>, dim3(n), dim3(kThreads), 0, st, I, W, m, h, terms, out);
    return post();
}

template <typename T>
hipError_t launch_delta_e(hipStream_t st, const Inst *I, const Work *W, int n, const T *m,
                          const T *h, uint32_t stage, uint32_t sweep, double *out)
{
    if (n <= 0) return hipSuccess;
    MMC_LAUNCH(k_delta_e<T>, dim3(n), dim3(kThreads), 0, st, I, W, m, h, stage, sweep, out);
    return post();
}

template <typename T>
hipError_t launch_init(hipStream_t st, const Inst *I, const Work *W, int n, T *m)
{
    if (n <= 0) return hipSuccess;
    MMC_LAUNCH(k_init<T>, dim3(n), dim3(kThreads), 0, st, I, W, m);
    return post();
}

hipError_t launch_set_ctl(hipStream_t st, uint32_t *ctl, uint32_t stage, uint32_t sweep_base)
{
    MMC_LAUNCH(k_set_ctl, dim3(1), dim3(64), 0, st, ctl, stage, sweep_base);
    return post();
}

template <typename T>
hipError_t set_lds_metro(int bytes)
{
    hipError_t e = hipSuccess;
    set_lds_one(e, (const void *)k_metro_tile<T>, bytes);
    return e;
}

template <typename T>
static hipError_t set_lds_t(int bytes)
{
    for (auto f : {set_lds_xpass<T>, set_lds_ypass<T>, set_lds_fused_inv<T>, set_lds_fused_fwd<T>,
                   set_lds_rows<T>, set_lds_metro<T>}) {
        const hipError_t e = f(bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t set_lds_limit(size_t bytes)
{
    hipError_t e = set_lds_t<float>((int)bytes);
    if (e != hipSuccess) return e;
    return set_lds_t<double>((int)bytes);
}

#define MMC_INST(T)                                                                           \
    template hipError_t launch_metro<T>(hipStream_t, const Inst *, const Work *, int, T *,    \
                                        const T *, const double *, const uint32_t *,          \
                                        uint32_t, int, int);                                  \
    template hipError_t launch_metro_tile<T>(hipStream_t, const Inst *, const Work *, int,    \
                                             T *, const T *, const double *, const uint32_t *, \
                                             uint32_t, int, size_t);                          \
    template hipError_t launch_energy<T>(hipStream_t, const Inst *, const Work *, int,        \
                                         const T *, const T *, double *, int);                \
    template hipError_t launch_field<T>(hipStream_t, const Inst *, const Work *, int,         \
                                        const T *, const T *, int, double *);                 \
    template hipError_t launch_energy_density<T>(hipStream_t, const Inst *, const Work *, int, \
                                                 const T *, const T *, int, double *);       \
    template hipError_t launch_delta_e<T>(hipStream_t, const Inst *, const Work *, int,       \
                                          const T *, const T *, uint32_t, uint32_t, double *); \
    template hipError_t launch_init<T>(hipStream_t, const Inst *, const Work *, int, T *);
MMC_INST(float)
MMC_INST(double)
#undef MMC_INST

}  // namespace mmc
