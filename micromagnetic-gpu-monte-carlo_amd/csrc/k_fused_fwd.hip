// k_fused_fwd.hip -- second checkerboard colour of the Metropolis update
// fused with the forward x pass of the updated rows.
#include "mmc_device.h"

namespace mmc {

// Second checkerboard colour of a tile of real rows fused with the forward x
// pass of the rows' final state (items and LDS rows as k_xinv_metro).  The
// rows are loaded into LDS packed two real rows per complex row, colour-1
// cells are updated from the compact h and written to global m and to LDS,
// then the rows are transformed and their half spectrum stored as k_xfwd_n
// does (zf = 0), or transformed along z in registers and stored to the
// xz-spectrum B as k_zfwd does (zf = 1; compiled only into the ZF variants,
// as in k_xinv_metro).
template <typename T, int N, int RZ, bool ZF>
__device__ __forceinline__ void metro_xfwd_tile(const Inst *__restrict__ I,
                                                         const Work w,
                                                         T *__restrict__ mb,
                                                         const T *__restrict__ hb,
                                                         cpx<T> *__restrict__ specb,
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
    const long NN = in.N;
    T *m = mb + in.m_off;
    const auto twr = load_twiddles<T, N>(twb + in.twx);
    // Colour-1 cells of this thread: e = threadIdx.x + q*kThreads < ncell
    // (at most QC).  Their compact h, neighbours (colour 0, not written by
    // this kernel) and trial moves are fetched CH cells at a time, loads
    // issued together; when they fit (PFC = QC) they are fetched together
    // with the tile's rows.
    const T *hc = hb + in.h_off;
    const int hw = (nx + 1) >> 1;
    const long cst = (long)hw * ny * nz;
    const uint32_t sweep = ctl[1] + sweep_off, stage = ctl[0];
    const int plane = nyt * hw;
    const int ncell = nzl * plane;
    constexpr int QC = (RZ * N + kThreads - 1) / kThreads;
    constexpr int PFC = prefetch_cells<T, QC>();
    constexpr int CH = PFC ? PFC : (QC < gather_chunk<T>() ? QC : gather_chunk<T>());
    const Div dhw(hw);
    auto cell_at = [&](int e, int &zl, int &yy, int &j) {
        zl = ZF ? e / plane : 0;   // one layer per tile unless z is folded in
        const int r = e - zl * plane;
        yy = dhw(r);
        j = r - yy * hw;
    };
    V3<T> chi[CH];
    Nbrs<T> cnb[CH];
    Move<T> cmv[CH];
    // Prefetching tiles of one layer (PFC, not folding z) load only the
    // inputs that are not in the tile's rows: the rows y - 1 and y + 1
    // (used when the cell's row is the tile's first or last, loaded either
    // way so that no load sits under a lane condition) and the sum of the z
    // neighbours (the one form the bond sums use); x neighbours and rows
    // inside the tile are read from the rows in LDS (colour-0 cells, which
    // this pass does not write).  Missing neighbours are zero, as in
    // gather_nbrs.
    const bool lrow = PFC > 0 && !zf;
    V3<T> pym[CH], pyp[CH], pzs[CH];
    const uint32_t unx = (uint32_t)nx, uny = (uint32_t)ny;
    auto keep = [](bool ok, const V3<T> &v) -> V3<T> {
        return {ok ? v.x : T(0), ok ? v.y : T(0), ok ? v.z : T(0)};
    };
    auto fetch = [&](int q, int k) {
        const int e = threadIdx.x + q * kThreads;
        int zl, yy, j;
        cell_at(e, zl, yy, j);
        const int z = z0 + zl, y = y0 + yy;
        const int x = 2 * j + ((y + z + 1) & 1);
        if (e < ncell && x < nx) {
            const long oh = ((long)z * ny + y) * hw + j;
            chi[k] = {hc[oh], hc[cst + oh], hc[2 * cst + oh]};
            if (lrow) {
                const int jm = nbr(y, -1, ny, in.pbc[1]), jp = nbr(y, +1, ny, in.pbc[1]);
                pym[k] = keep(jm >= 0, load_m(m, (uint32_t)NN, cell_index(unx, uny, x, jm >= 0 ? jm : y, z)));
                pyp[k] = keep(jp >= 0, load_m(m, (uint32_t)NN, cell_index(unx, uny, x, jp >= 0 ? jp : y, z)));
                V3<T> zp = {T(0), T(0), T(0)}, zm = zp;
                int jz;
                if (nz > 1 && (jz = nbr(z, +1, nz, in.pbc[2])) >= 0)
                    zp = load_m(m, (uint32_t)NN, cell_index(unx, uny, x, y, jz));
                if (nz > 1 && (jz = nbr(z, -1, nz, in.pbc[2])) >= 0)
                    zm = load_m(m, (uint32_t)NN, cell_index(unx, uny, x, y, jz));
                pzs[k] = {zp.x + zm.x, zp.y + zm.y, zp.z + zm.z};
            } else {
                cnb[k] = gather_nbrs(in, m, x, y, z);
            }
            cmv[k] = trial_move<T>(in, (uint32_t)((long)nx * (y + (long)ny * z) + x), sweep,
                                   stage);
        }
    };
    auto row_ld = [&](int e) {
        const int c = e / (RZ * N), r = e - c * (RZ * N);
        const int zr = r / N, x = r - zr * N;
        const int zl = zr / R, q = zr - zl * R;
        T re = T(0), im = T(0);
        if (x < nx && 2 * q < nyt) {
            const T *row = m + c * NN + (long)nx * ((y0 + 2 * q) + (long)ny * (z0 + zl));
            re = row[x];
            if (2 * q + 1 < nyt) im = row[(long)nx + x];
        }
        return cpx<T>{re, im};
    };
    auto row_st = [&](int e, cpx<T> v) { s[e] = v; };
    // Tiles of one complete row pair of one layer with more than two cells
    // per thread (LNB) issue, with the row loads, every colour-1 input that
    // is not in the tile's rows: the compact h, the neighbour row outside
    // the pair, the z neighbours (as their sum, the only form the bond sums
    // use); the x neighbours and the pair's other row are then read from the
    // rows in LDS (colour-0 cells, which this pass does not write), so the
    // Metropolis phase issues no global load.  Missing neighbours are zero,
    // as in gather_nbrs.
    constexpr bool LNB = PFC == 0 && RZ == 1 && QC > 2 && QC <= 4;
    const bool lnb = LNB && !zf && nyt == 2;
    V3<T> ph[LNB ? QC : 1], po[LNB ? QC : 1], pz[LNB ? QC : 1];
    if (LNB && lnb) {
        const auto rows = staged_load<3 * RZ * N, kThreads, cpx<T>>(3 * RZ * N, row_ld);
        const int jzp = nz > 1 ? nbr(z0, +1, nz, in.pbc[2]) : -1;
        const int jzm = nz > 1 ? nbr(z0, -1, nz, in.pbc[2]) : -1;
        #pragma unroll
        for (int q = 0; q < (LNB ? QC : 0); ++q) {
            const int e = threadIdx.x + q * kThreads;
            int zl, yy, j;
            cell_at(e, zl, yy, j);
            const int y = y0 + yy;
            const int x = 2 * j + ((y + z0 + 1) & 1);
            // a slot without a cell reads cell (0, y) and its values go unused
            const int xs = (e < ncell && x < nx) ? x : 0;
            const long oh = ((long)z0 * ny + y) * hw + j;
            ph[q] = {hc[oh], hc[cst + oh], hc[2 * cst + oh]};
            const int jo = nbr(y, yy ? +1 : -1, ny, in.pbc[1]);
            po[q] = keep(jo >= 0, load_m(m, (uint32_t)NN, cell_index(unx, uny, xs, jo >= 0 ? jo : y, z0)));
            V3<T> zp = {T(0), T(0), T(0)}, zm = zp;
            if (jzp >= 0) zp = load_m(m, (uint32_t)NN, cell_index(unx, uny, xs, y, jzp));
            if (jzm >= 0) zm = load_m(m, (uint32_t)NN, cell_index(unx, uny, xs, y, jzm));
            pz[q] = {zp.x + zm.x, zp.y + zm.y, zp.z + zm.z};
        }
        staged_store<3 * RZ * N, kThreads, cpx<T>>(3 * RZ * N, rows, row_st);
    } else if (PFC) {
        const auto rows = staged_load<3 * RZ * N, kThreads, cpx<T>>(3 * RZ * N, row_ld);
        #pragma unroll
        for (int q = 0; q < PFC; ++q) fetch(q, q);
        staged_store<3 * RZ * N, kThreads, cpx<T>>(3 * RZ * N, rows, row_st);
    } else {
        staged_copy<3 * RZ * N, kThreads, cpx<T>>(3 * RZ * N, row_ld, row_st);
    }
    store_twiddles<T, N>(tw, twr);
    __syncthreads();
    if (LNB && lnb) {
        const T bt = T(beta[w.inst]);
        auto lds_m = [&](int yy, int x) -> V3<T> {
            const cpx<T> a = s[x], b = s[N + x], c = s[2 * N + x];
            return yy ? V3<T>{a.y, b.y, c.y} : V3<T>{a.x, b.x, c.x};
        };
        #pragma unroll
        for (int q = 0; q < (LNB ? QC : 0); ++q) {
            const int e = threadIdx.x + q * kThreads;
            if (e >= ncell) break;
            int zl, yy, j;
            cell_at(e, zl, yy, j);
            const int y = y0 + yy;
            const int x = 2 * j + ((y + z0 + 1) & 1);
            if (x >= nx) continue;
            Nbrs<T> nb;
            const int jp = nbr(x, +1, nx, in.pbc[0]), jm = nbr(x, -1, nx, in.pbc[0]);
            nb.xp = keep(jp >= 0, lds_m(yy, jp >= 0 ? jp : x));
            nb.xm = keep(jm >= 0, lds_m(yy, jm >= 0 ? jm : x));
            const V3<T> vin = lds_m(yy ^ 1, x);
            nb.yp = yy ? po[q] : vin;
            nb.ym = yy ? vin : po[q];
            nb.zp = pz[q];
            nb.zm = {T(0), T(0), T(0)};
            const V3<T> mi = lds_m(yy, x);
            const long i = (long)nx * (y + (long)ny * z0) + x;
            const Move<T> mv = trial_move<T>(in, (uint32_t)i, sweep, stage);
            const V3<T> mn = metro_apply(in, m, i, mi, nb, ph[q], mv, bt);
            if (yy) {
                s[x].y = mn.x;
                s[N + x].y = mn.y;
                s[2 * N + x].y = mn.z;
            } else {
                s[x].x = mn.x;
                s[N + x].x = mn.y;
                s[2 * N + x].x = mn.z;
            }
        }
    } else {
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
                const int x = 2 * j + ((y + z + 1) & 1);
                if (x >= nx) continue;
                const int o = (zl * R + (yy >> 1)) * N + x;
                const bool odd = yy & 1;
                const V3<T> mi = odd ? V3<T>{s[o].y, s[RZ * N + o].y, s[2 * RZ * N + o].y}
                                     : V3<T>{s[o].x, s[RZ * N + o].x, s[2 * RZ * N + o].x};
                const long i = (long)nx * (y + (long)ny * z) + x;
                if (lrow) {
                    // tile row r (layer 0), column c
                    auto row_m = [&](int r, int c) -> V3<T> {
                        const int u = (r >> 1) * N + c;
                        return (r & 1) ? V3<T>{s[u].y, s[RZ * N + u].y, s[2 * RZ * N + u].y}
                                       : V3<T>{s[u].x, s[RZ * N + u].x, s[2 * RZ * N + u].x};
                    };
                    const int jp = nbr(x, +1, nx, in.pbc[0]), jm = nbr(x, -1, nx, in.pbc[0]);
                    cnb[k].xp = keep(jp >= 0, row_m(yy, jp >= 0 ? jp : x));
                    cnb[k].xm = keep(jm >= 0, row_m(yy, jm >= 0 ? jm : x));
                    const V3<T> vm = row_m(yy > 0 ? yy - 1 : 0, x), vp = row_m(yy + 1 < nyt ? yy + 1 : yy, x);
                    cnb[k].ym = yy > 0 ? vm : pym[k];
                    cnb[k].yp = yy + 1 < nyt ? vp : pyp[k];
                    cnb[k].zp = pzs[k];
                    cnb[k].zm = {T(0), T(0), T(0)};
                }
                const V3<T> mn = metro_apply(in, m, i, mi, cnb[k], chi[k], cmv[k], bt);
                if (odd) {
                    s[o].y = mn.x;
                    s[RZ * N + o].y = mn.y;
                    s[2 * RZ * N + o].y = mn.z;
                } else {
                    s[o].x = mn.x;
                    s[RZ * N + o].x = mn.y;
                    s[2 * RZ * N + o].x = mn.z;
                }
            }
        }
    }
    __syncthreads();
    fft_fixed<T, N, 3 * RZ>(s, tw);
    const int npair = (nyt + 1) >> 1;
    const int per = H * npair;
    const Div dper(per), dnp(npair);
    if (!zf) {
        const long cstride = (long)H * nz * ny;
        cpx<T> *dst = specb + in.a_off;
        for (int e = threadIdx.x; e < 3 * per; e += kThreads) {
            const int c = dper(e), r = e - c * per;
            const int k = dnp(r), q = r - k * npair;
            const cpx<T> *row = s + (c * RZ + q) * N;
            const cpx<T> A = row[k];
            const cpx<T> B = row[k ? N - k : 0];
            cpx<T> *o = dst + c * cstride + ((long)k * nz + z0) * ny + y0 + 2 * q;
            const cpx<T> oa = {(A.x + B.x) * T(0.5), (A.y - B.y) * T(0.5)};
            const cpx<T> ob = {(A.y + B.y) * T(0.5), (B.x - A.x) * T(0.5)};
            if (p16) {
                store_pair16(o, oa, ob);
            } else {
                o[0] = oa;
                if (2 * q + 1 < nyt) o[1] = ob;
            }
        }
    } else {
        const int Pz = in.Pz;
        const long cstride = (long)H * Pz * ny;
        cpx<T> *dst = specb + in.b_off;
        const cpx<T> *twz = twb + in.twz;
        for (int e = threadIdx.x; e < 3 * per; e += kThreads) {
            const int c = dper(e), r = e - c * per;
            const int k = dnp(r), q = r - k * npair;
            const bool hasb = 2 * q + 1 < nyt;
            cpx<T> xa[kZDirect], xb[kZDirect], oa[kZDirect], ob[kZDirect];
            #pragma unroll
            for (int zl = 0; zl < kZDirect; ++zl) {
                xa[zl] = xb[zl] = {T(0), T(0)};
                if (zl < nz) {
                    const cpx<T> *row = s + (c * RZ + zl * R + q) * N;
                    const cpx<T> A = row[k];
                    const cpx<T> B = row[k ? N - k : 0];
                    xa[zl] = {(A.x + B.x) * T(0.5), (A.y - B.y) * T(0.5)};
                    xb[zl] = {(A.y + B.y) * T(0.5), (B.x - A.x) * T(0.5)};
                }
            }
            zdft_regs<T, false>(xa, oa, nz, Pz, Pz, twz);
            zdft_regs<T, false>(xb, ob, nz, Pz, Pz, twz);
            cpx<T> *o = dst + c * cstride + (long)k * Pz * ny + y0 + 2 * q;
            #pragma unroll
            for (int kz = 0; kz < kZDirect; ++kz) {
                if (kz < Pz) {
                    o[(long)kz * ny] = oa[kz];
                    if (hasb) o[(long)kz * ny + 1] = ob[kz];
                }
            }
        }
    }
}

template <typename T, int N, int RZ, bool ZF>
__global__ __launch_bounds__(kThreads) void k_metro_xfwd(const Inst *__restrict__ I,
                                                         const Work *__restrict__ W,
                                                         T *__restrict__ mb,
                                                         const T *__restrict__ hb,
                                                         cpx<T> *__restrict__ specb,
                                                         const double *__restrict__ beta,
                                                         const uint32_t *__restrict__ ctl,
                                                         uint32_t sweep_off,
                                                         const cpx<T> *__restrict__ twb)
{
    metro_xfwd_tile<T, N, RZ, ZF>(I, W[blockIdx.x], mb, hb, specb, beta, ctl, sweep_off, twb);
}

// One launch over items of every x length of a batch (any_x_len), tile
// policy POL (any_tile).
template <typename T, int POL>
__global__ __launch_bounds__(kThreads) void k_metro_xfwd_any(const Inst *__restrict__ I,
                                                             const Work *__restrict__ W,
                                                             T *__restrict__ mb,
                                                             const T *__restrict__ hb,
                                                             cpx<T> *__restrict__ specb,
                                                             const double *__restrict__ beta,
                                                             const uint32_t *__restrict__ ctl,
                                                             uint32_t sweep_off,
                                                             const cpx<T> *__restrict__ twb)
{
    const Work w = W[blockIdx.x];
    switch (I[w.inst].Px) {
#define MMC_ANY(NN)                                                                              \
    case NN:                                                                                     \
        metro_xfwd_tile<T, NN, any_tile(POL, NN), false>(I, w, mb, hb, specb, beta, ctl, sweep_off, \
                                                         twb);                                   \
        break;
    MMC_ANY(128) MMC_ANY(192) MMC_ANY(256) MMC_ANY(384) MMC_ANY(512)
#undef MMC_ANY
    }
}

template <typename T>
hipError_t launch_metro_xfwd(hipStream_t st, const Inst *I, const Work *W, int n, int fixed_n,
                             int rows, bool zf, T *m, const T *h, cpx<T> *spec, const double *beta,
                             const uint32_t *ctl, uint32_t sweep_off, const cpx<T> *tw, size_t lds)
{
    if (n <= 0) return hipSuccess;
    const dim3 g(n), b(kThreads);
    constexpr int F = fused_rows_base<T>();
    if (fixed_n == kAnyLen) {
        switch (rows) {
        case 0: MMC_LAUNCH((k_metro_xfwd_any<T, 0>), g, b, lds, st, I, W, m, h, spec, beta, ctl, sweep_off, tw); break;
        case 1: MMC_LAUNCH((k_metro_xfwd_any<T, 1>), g, b, lds, st, I, W, m, h, spec, beta, ctl, sweep_off, tw); break;
        case 2: MMC_LAUNCH((k_metro_xfwd_any<T, 2>), g, b, lds, st, I, W, m, h, spec, beta, ctl, sweep_off, tw); break;
        default: MMC_LAUNCH((k_metro_xfwd_any<T, 3>), g, b, lds, st, I, W, m, h, spec, beta, ctl, sweep_off, tw); break;
        }
        return post();
    }
#define MMC_MX_CASE(TT, NN, RR)                                                                  \
    if (fixed_n == NN && rows == RR) {                                                           \
        if (zf)                                                                                  \
            MMC_LAUNCH((k_metro_xfwd<TT, NN, RR, true>), g, b, lds, st, I, W, m, h, spec, beta,  \
                       ctl, sweep_off, tw);                                                      \
        else                                                                                     \
            MMC_LAUNCH((k_metro_xfwd<TT, NN, RR, false>), g, b, lds, st, I, W, m, h, spec, beta, \
                       ctl, sweep_off, tw);                                                      \
        return post();                                                                           \
    }
    MMC_FUSED(MMC_MX_CASE, T, F)
#undef MMC_MX_CASE
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t set_lds_fused_fwd(int bytes)
{
    hipError_t e = hipSuccess;
    constexpr int F = fused_rows_base<T>();
#define MMC_FL(TT, NN, RR) set_lds_one(e, (const void *)k_metro_xfwd<TT, NN, RR, false>, bytes); \
    set_lds_one(e, (const void *)k_metro_xfwd<TT, NN, RR, true>, bytes);
    MMC_FUSED(MMC_FL, T, F)
#undef MMC_FL
    set_lds_one(e, (const void *)k_metro_xfwd_any<T, 0>, bytes);
    set_lds_one(e, (const void *)k_metro_xfwd_any<T, 1>, bytes);
    set_lds_one(e, (const void *)k_metro_xfwd_any<T, 2>, bytes);
    set_lds_one(e, (const void *)k_metro_xfwd_any<T, 3>, bytes);
    return e;
}

#define MMC_INST(T)                                                                           \
    template hipError_t launch_metro_xfwd<T>(hipStream_t, const Inst *, const Work *, int, int, \
                                             int, bool, T *, const T *, cpx<T> *, const double *, \
                                             const uint32_t *, uint32_t, const cpx<T> *,     \
                                             size_t);                                        \
    template hipError_t set_lds_fused_fwd<T>(int);
MMC_INST(float)
MMC_INST(double)
#undef MMC_INST

}  // namespace mmc
