// k_fused_inv.hip -- inverse x pass fused with the first checkerboard
// colour of the Metropolis update (mcKernel2, src/mc_sampling.h:27-41).
#include "mmc_device.h"

namespace mmc {

// Inverse x pass of all three components of a tile of real rows fused with
// the first checkerboard colour of those rows.  Item (inst, 0, z0, y0): RZ
// complex rows per component, each two real rows packed.  An instance with
// zf = 0 covers layer z0, rows y0.. (R = RZ row pairs) and reads the
// x-spectrum A; with zf = 1 (short z, Pz <= kZDirect) it covers all nz
// layers (R = RZ / nz row pairs per layer) and reads the xz-spectrum B,
// inverting z in registers first.  LDS row (comp, layer zl, pair q) =
// comp*RZ + zl*R + q.  The h rows stay in LDS for colour 0 (cells with
// x + y + z even); the h of colour-1 cells is stored compactly for the
// second pass: hc[comp][(z*ny + y)*hw + j], x = 2j + ((y + z + 1) & 1),
// hw = ceil(nx/2), in the instance's h region.  Same arithmetic as
// k_zinv + k_xinv_n + k_metro.  Tiles that fold z in (zf) are compiled only
// into the ZF variants: their unrolled z transform's bounds otherwise cost
// every tile of the common kernels scalar registers and spill code.
template <typename T, int N, int RZ, bool ZF>
__device__ __forceinline__ void xinv_metro_tile(const Inst *__restrict__ I,
                                                         const Work w,
                                                         const cpx<T> *__restrict__ specb,
                                                         T *__restrict__ mb, T *__restrict__ hb,
                                                         const double *__restrict__ beta,
                                                         const uint32_t *__restrict__ ctl,
                                                         uint32_t sweep_off,
                                                         const cpx<T> *__restrict__ twb)
{
    cpx<T> *s = reinterpret_cast<cpx<T> *>(g_smem);
    cpx<T> *tw = s + 3 * RZ * N;
    const Inst &in = I[w.inst];
    const bool p16 = pair16<T>(in);
    const int z0 = w.b, y0 = w.c;
    constexpr int H = N / 2 + 1;
    const int nx = in.nx, ny = in.ny, nz = in.nz;
    const bool zf = ZF && in.zf;
    const int nzl = zf ? nz : 1;
    const int R = RZ / nzl;
    const int nyt = min(2 * R, ny - y0);
    const int npair = (nyt + 1) >> 1;
    const auto twr = load_twiddles<T, N>(twb + in.twx);
    // Colour-0 cells of this thread: e = threadIdx.x + q*kThreads < ncell
    // (at most QC).  Their m values, neighbours and trial moves are fetched
    // after the transform, CH cells at a time with the loads issued together
    // (holding them across the transform costs more occupancy than it saves
    // latency).  Nothing in this kernel writes a cell they read: colour-0
    // cells only have colour-1 neighbours.
    T *m = mb + in.m_off;
    const long NN = in.N;
    const int hw = (nx + 1) >> 1;
    const int plane = nyt * hw;
    const int ncell = nzl * plane;
    const uint32_t sweep = ctl[1] + sweep_off, stage = ctl[0];
    constexpr int QC = (RZ * N + kThreads - 1) / kThreads;
    constexpr int PFC = 0;
    constexpr int CH = QC < gather_chunk<T>() ? QC : gather_chunk<T>();
    const Div dhw(hw);
    auto cell_at = [&](int e, int &zl, int &yy, int &j) {
        zl = ZF ? e / plane : 0;   // one layer per tile unless z is folded in
        const int r = e - zl * plane;
        yy = dhw(r);
        j = r - yy * hw;
    };
    V3<T> cmi[CH];
    Nbrs<T> cnb[CH];
    // Trial moves depend on the cell and the counters only: tiles whose
    // spectrum comes from A (or B, pair-major) draw all of theirs while the
    // spectrum loads are in flight, when the vector ALU would otherwise
    // wait, and keep them across the transform; tiles that fold z in draw
    // each with the cell's other inputs.
    const bool em = !zf;
    Move<T> cmv[QC];
    auto cell_id = [&](int e, int &xa, int &y, int &z) {
        int zl, yy, j;
        cell_at(e, zl, yy, j);
        z = z0 + zl;
        y = y0 + yy;
        xa = 2 * j + ((y + z) & 1);
    };
    auto draw_all = [&] {
        #pragma unroll
        for (int q = 0; q < QC; ++q) {
            const int e = threadIdx.x + q * kThreads;
            int xa, y, z;
            cell_id(e, xa, y, z);
            if (e < ncell && xa < nx)
                cmv[q] = trial_move<T>(in, (uint32_t)((long)nx * (y + (long)ny * z) + xa), sweep, stage);
        }
    };
    auto fetch = [&](int q, int k) {
        const int e = threadIdx.x + q * kThreads;
        int xa, y, z;
        cell_id(e, xa, y, z);
        if (e < ncell && xa < nx) {
            const long i = (long)nx * (y + (long)ny * z) + xa;
            cmi[k] = load_m(m, NN, i);
            cnb[k] = gather_nbrs(in, m, xa, y, z);
            if (!em) cmv[q] = trial_move<T>(in, (uint32_t)i, sweep, stage);
        }
    };
    if (npair < R) {
        const int zr = (R - npair) * N;
        for (int e = threadIdx.x; e < 3 * nzl * zr; e += kThreads) {
            const int cl = e / zr, r = e - cl * zr;
            const int c = cl / nzl, zl = cl - c * nzl;
            s[(c * RZ + zl * R + npair) * N + r] = {T(0), T(0)};
        }
    }
    const int per = H * npair;
    const Div dper(per), dnp(npair);
    if (!zf) {
        // the spectrum as the y pass left it: y-contiguous in A, or (Inst::a2)
        // pair-major at pm_off, [3][nz][ny/2][Hx][2], where a row pair's
        // spectrum is contiguous (whole lines instead of 16 bytes of each)
        const bool pm = in.a2;
        const long cstride = (long)H * nz * ny;
        const cpx<T> *src = specb + (pm ? in.pm_off : in.a_off);
        const int np2 = ny >> 1;
        const bool p16x = pm ? sizeof(T) == 4 : p16;
        // element e -> (component c, frequency k, pair q): pairs innermost
        // for the y-contiguous layout, frequencies innermost for pair-major
        auto at = [&](int e, int &c, int &k, int &q) {
            c = dper(e);
            const int r = e - c * per;
            const int qa = r / H, qb = dnp(r);
            q = pm ? qa : r - qb * npair;
            k = pm ? r - qa * H : qb;
        };
        staged_copy<3 * H * RZ, kThreads, Pair2<cpx<T>>>(
            3 * per,
            [&](int e) {
                int c, k, q;
                at(e, c, k, q);
                const cpx<T> *p =
                    pm ? src + ((((long)c * nz + z0) * np2 + (y0 >> 1) + q) * H + k) * 2
                       : src + c * cstride + ((long)k * nz + z0) * ny + y0 + 2 * q;
                if (p16x) return load_pair16(p);
                Pair2<cpx<T>> v;
                // branch-free: p[1] read in place of a missing row, zero selected
                const bool hb = 2 * q + 1 < nyt;
                v.a = p[0];
                const cpx<T> b = p[hb ? 1 : 0];
                v.b = {hb ? b.x : T(0), hb ? b.y : T(0)};
                return v;
            },
            [&](int e, Pair2<cpx<T>> v) {
                int c, k, q;
                at(e, c, k, q);
                const cpx<T> a = v.a;
                const T bx = v.b.x, by = v.b.y;
                cpx<T> *row = s + (c * RZ + q) * N;
                row[k] = {a.x - by, -(a.y + bx)};
                if (k > 0 && N - k >= H) row[N - k] = {a.x + by, a.y - bx};
            },
            draw_all);
    } else {
        const int Pz = in.Pz;
        const long cstride = (long)H * Pz * ny;
        const cpx<T> *src = specb + in.b_off;
        const cpx<T> *twz = twb + in.twz;
        for (int e = threadIdx.x; e < 3 * per; e += kThreads) {
            const int c = e / per, r = e - c * per;
            const int k = r / npair, q = r - k * npair;
            const cpx<T> *p = src + c * cstride + (long)k * Pz * ny + y0 + 2 * q;
            const bool hasb = 2 * q + 1 < nyt;
            cpx<T> va[kZDirect], vb[kZDirect], oa[kZDirect], ob[kZDirect];
            #pragma unroll
            for (int kz = 0; kz < kZDirect; ++kz) {
                va[kz] = kz < Pz ? p[(long)kz * ny] : cpx<T>{T(0), T(0)};
                vb[kz] = (kz < Pz && hasb) ? p[(long)kz * ny + 1] : cpx<T>{T(0), T(0)};
            }
            zdft_regs<T, true>(va, oa, Pz, nz, Pz, twz);
            zdft_regs<T, true>(vb, ob, Pz, nz, Pz, twz);
            #pragma unroll
            for (int zl = 0; zl < kZDirect; ++zl) {
                if (zl < nz) {
                    const cpx<T> a = oa[zl];
                    const T bx = hasb ? ob[zl].x : T(0), by = hasb ? ob[zl].y : T(0);
                    cpx<T> *row = s + (c * RZ + zl * R + q) * N;
                    row[k] = {a.x - by, -(a.y + bx)};
                    if (k > 0 && N - k >= H) row[N - k] = {a.x + by, a.y - bx};
                }
            }
        }
    }
    store_twiddles<T, N>(tw, twr);
    __syncthreads();
    fft_fixed<T, N, 3 * RZ>(s, tw);
    auto hval = [&](int zl, int yy, int x) -> V3<T> {
        const int o = (zl * R + (yy >> 1)) * N + x;
        if (yy & 1) return {-s[o].y, -s[RZ * N + o].y, -s[2 * RZ * N + o].y};
        return {s[o].x, s[RZ * N + o].x, s[2 * RZ * N + o].x};
    };
    T *hc = hb + in.h_off;
    const long cst = (long)hw * ny * nz;
    const T bt = T(beta[w.inst]);
    #pragma unroll
    for (int q0 = 0; q0 < QC; q0 += CH) {
        if (!PFC) {
            #pragma unroll
            for (int k = 0; k < CH; ++k)
                if (q0 + k < QC) fetch(q0 + k, k);
        }
        #pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int e = threadIdx.x + (q0 + k) * kThreads;
            if (q0 + k >= QC || e >= ncell) break;
            int zl, yy, j;
            cell_at(e, zl, yy, j);
            const int z = z0 + zl, y = y0 + yy;
            const int p0 = (y + z) & 1;
            const int xa = 2 * j + p0, xb = 2 * j + (p0 ^ 1);
            const long rowi = (long)nx * (y + (long)ny * z);
            if (xb < nx) {
                const V3<T> h1 = hval(zl, yy, xb);
                const long o = ((long)z * ny + y) * hw + j;
                hc[o] = h1.x;
                hc[cst + o] = h1.y;
                hc[2 * cst + o] = h1.z;
            }
            if (xa < nx)
                metro_apply(in, m, rowi + xa, cmi[k], cnb[k], hval(zl, yy, xa), cmv[q0 + k], bt);
        }
    }
}

template <typename T, int N, int RZ, bool ZF>
__global__ __launch_bounds__(kThreads) void k_xinv_metro(const Inst *__restrict__ I,
                                                         const Work *__restrict__ W,
                                                         const cpx<T> *__restrict__ specb,
                                                         T *__restrict__ mb, T *__restrict__ hb,
                                                         const double *__restrict__ beta,
                                                         const uint32_t *__restrict__ ctl,
                                                         uint32_t sweep_off,
                                                         const cpx<T> *__restrict__ twb)
{
    xinv_metro_tile<T, N, RZ, ZF>(I, W[blockIdx.x], specb, mb, hb, beta, ctl, sweep_off, twb);
}

// One launch over items of every x length of a batch (any_x_len), tile
// policy POL (any_tile).
template <typename T, int POL>
__global__ __launch_bounds__(kThreads) void k_xinv_metro_any(const Inst *__restrict__ I,
                                                             const Work *__restrict__ W,
                                                             const cpx<T> *__restrict__ specb,
                                                             T *__restrict__ mb, T *__restrict__ hb,
                                                             const double *__restrict__ beta,
                                                             const uint32_t *__restrict__ ctl,
                                                             uint32_t sweep_off,
                                                             const cpx<T> *__restrict__ twb)
{
    const Work w = W[blockIdx.x];
    switch (I[w.inst].Px) {
#define MMC_ANY(NN)                                                                              \
    case NN:                                                                                     \
        xinv_metro_tile<T, NN, any_tile(POL, NN), false>(I, w, specb, mb, hb, beta, ctl, sweep_off, \
                                                         twb);                                   \
        break;
    MMC_ANY(128) MMC_ANY(192) MMC_ANY(256) MMC_ANY(384) MMC_ANY(512)
#undef MMC_ANY
    }
}

template <typename T>
hipError_t launch_xinv_metro(hipStream_t st, const Inst *I, const Work *W, int n, int fixed_n,
                             int rows, bool zf, const cpx<T> *spec, T *m, T *h, const double *beta,
                             const uint32_t *ctl, uint32_t sweep_off, const cpx<T> *tw, size_t lds)
{
    if (n <= 0) return hipSuccess;
    const dim3 g(n), b(kThreads);
    constexpr int F = fused_rows_base<T>();
    if (fixed_n == kAnyLen) {
        switch (rows) {
        case 0: MMC_LAUNCH((k_xinv_metro_any<T, 0>), g, b, lds, st, I, W, spec, m, h, beta, ctl, sweep_off, tw); break;
        case 1: MMC_LAUNCH((k_xinv_metro_any<T, 1>), g, b, lds, st, I, W, spec, m, h, beta, ctl, sweep_off, tw); break;
        case 2: MMC_LAUNCH((k_xinv_metro_any<T, 2>), g, b, lds, st, I, W, spec, m, h, beta, ctl, sweep_off, tw); break;
        default: MMC_LAUNCH((k_xinv_metro_any<T, 3>), g, b, lds, st, I, W, spec, m, h, beta, ctl, sweep_off, tw); break;
        }
        return post();
    }
#define MMC_XM_CASE(TT, NN, RR)                                                                  \
    if (fixed_n == NN && rows == RR) {                                                           \
        if (zf)                                                                                  \
            MMC_LAUNCH((k_xinv_metro<TT, NN, RR, true>), g, b, lds, st, I, W, spec, m, h, beta,  \
                       ctl, sweep_off, tw);                                                      \
        else                                                                                     \
            MMC_LAUNCH((k_xinv_metro<TT, NN, RR, false>), g, b, lds, st, I, W, spec, m, h, beta, \
                       ctl, sweep_off, tw);                                                      \
        return post();                                                                           \
    }
    MMC_FUSED(MMC_XM_CASE, T, F)
#undef MMC_XM_CASE
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t set_lds_fused_inv(int bytes)
{
    hipError_t e = hipSuccess;
    constexpr int F = fused_rows_base<T>();
#define MMC_FL(TT, NN, RR) set_lds_one(e, (const void *)k_xinv_metro<TT, NN, RR, false>, bytes); \
    set_lds_one(e, (const void *)k_xinv_metro<TT, NN, RR, true>, bytes);
    MMC_FUSED(MMC_FL, T, F)
#undef MMC_FL
    set_lds_one(e, (const void *)k_xinv_metro_any<T, 0>, bytes);
    set_lds_one(e, (const void *)k_xinv_metro_any<T, 1>, bytes);
    set_lds_one(e, (const void *)k_xinv_metro_any<T, 2>, bytes);
    set_lds_one(e, (const void *)k_xinv_metro_any<T, 3>, bytes);
    return e;
}

#define MMC_INST(T)                                                                           \
    template hipError_t launch_xinv_metro<T>(hipStream_t, const Inst *, const Work *, int, int, \
                                             int, bool, const cpx<T> *, T *, T *, const double *, \
                                             const uint32_t *, uint32_t, const cpx<T> *,     \
                                             size_t);                                        \
    template hipError_t set_lds_fused_inv<T>(int);
MMC_INST(float)
MMC_INST(double)
#undef MMC_INST

}  // namespace mmc
