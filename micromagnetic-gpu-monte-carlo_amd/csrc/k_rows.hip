// k_rows.hip -- whole Metropolis sweeps of row pairs of films of one or
// two layers: inverse x pass with halo pairs, both colours, forward x pass.
#include "mmc_device.h"

namespace mmc {

// One whole Metropolis sweep of a pair of real rows (y0, y0+1), in every one
// of the NZ layers, of a checkerboard film in one launch (with the y pass,
// two launches per sweep instead of three).  The tile inverts the x spectrum
// of its pair and of both neighbouring pairs in every layer (pairs start at
// even rows, as in their own tiles, so the arithmetic is theirs), updates
// colour 0 on rows y0-1..y0+2 (the halo rows' updates recomputed exactly as
// their own tiles compute them), colour 1 on its pair, and transforms the
// pair's final state forward into A.  Every read of m is of the pre-sweep
// state m_in, which no workgroup writes; the pair's final state goes to
// m_out.  LDS row ((p*NZ + z)*3 + comp) holds layer z of pair p = 0, 1, 2 <->
// rows y0-2.., y0.., y0+2.. (two real rows packed per complex row); after
// the inverse transform a colour-0 cell's slot is overwritten with its new
// value, then a colour-1 cell's.  Same arithmetic as k_xinv_metro +
// k_metro_xfwd with one row pair per layer.
template <typename T, int N, int NZ>
__global__ __launch_bounds__(kRowThreads) void k_row_sweep(const Inst *__restrict__ I,
                                                           const Work *__restrict__ W,
                                                           cpx<T> *__restrict__ specb,
                                                           const T *__restrict__ mib,
                                                           T *__restrict__ mob,
                                                           const double *__restrict__ beta,
                                                           const uint32_t *__restrict__ ctl,
                                                           uint32_t sweep_off,
                                                           const cpx<T> *__restrict__ twb)
{
    constexpr int H = N / 2 + 1;
    constexpr int NT = kRowThreads;
    constexpr int PR = 3 * NZ * 3;                 // LDS rows: pairs x layers x components
    cpx<T> *s = reinterpret_cast<cpx<T> *>(g_smem);
    cpx<T> *tw = s + PR * N;
    const Work w = W[blockIdx.x];
    const Inst &in = I[w.inst];
    const bool p16 = pair16<T>(in);
    const int y0 = w.c;
    const int nx = in.nx, ny = in.ny, pery = in.pbc[1];
    const int hw = (nx + 1) >> 1;
    const long NN = in.N;
    const T *mi_ = mib + in.m_off;
    T *mo = mob + in.m_off;
    cpx<T> *A = specb + in.a_off;
    const long cstride = (long)H * NZ * ny;     // component stride in A
    const uint32_t sweep = ctl[1] + sweep_off, stage = ctl[0];
    const T bt = T(beta[w.inst]);
    // absolute row of tile-relative row r (-2..3), or -1 outside an open axis
    auto row_of = [&](int r) {
        int y = y0 + r;
        if (y < 0 || y >= ny) {
            if (!pery) return -1;
            y = (y % ny + ny) % ny;
        }
        return y;
    };
    auto cell = [&](int y, int z, int x) { return cell_index(nx, ny, x, y, z); };
    const auto twr = load_twiddles<T, N, NT>(twb + in.twx);
    // Cells of this thread: colour 0 on rows -1..2 (QC0 at most, CH0 per
    // gather) and colour 1 on rows 0, 1 (QC1 at most), all layers.  Their
    // inputs from m_in and their trial moves depend on nothing this kernel
    // computes, so for short rows (PF) they are all fetched before the
    // transform, behind the spectrum loads: FP32 up to two colour-0 and one
    // colour-1 cell per thread; FP64, whose values take twice the registers,
    // single layers up to 384 points (one cell of each colour).
    constexpr int QC0 = (4 * NZ * (N / 2) + NT - 1) / NT;
    constexpr int QC1 = (2 * NZ * (N / 2) + NT - 1) / NT;
    constexpr bool PF = sizeof(T) == 4 ? QC0 <= 2 && QC1 <= 1 : NZ == 1 && N <= 384;
    constexpr int CH0 = PF ? QC0 : (QC0 < gather_chunk<T>() ? QC0 : gather_chunk<T>());
    constexpr int CH1 = PF ? QC1 : (QC1 < gather_chunk<T>() ? QC1 : gather_chunk<T>());
    const int ncell0 = 4 * NZ * hw, ncell1 = 2 * NZ * hw;
    // cell e = ((r - r0)*NZ + z)*hw + j; y < 0: no cell
    const Div dhw(hw);
    auto cell0 = [&](int q, int &r, int &z, int &y, int &x) {
        const int e = threadIdx.x + q * NT;
        const int rz = dhw(e), j = e - rz * hw;
        r = rz / NZ - 1;
        z = rz - (r + 1) * NZ;
        y = q < QC0 && e < ncell0 ? row_of(r) : -1;
        x = 2 * j + ((y + z) & 1);
        if (x >= nx) y = -1;
    };
    auto cell1 = [&](int q, int &r, int &z, int &y, int &x) {
        const int e = threadIdx.x + q * NT;
        const int rz = dhw(e), j = e - rz * hw;
        r = rz / NZ;
        z = rz - r * NZ;
        y = q < QC1 && e < ncell1 ? row_of(r) : -1;
        x = 2 * j + ((y + z + 1) & 1);
        if (x >= nx) y = -1;
    };
    V3<T> c0mi[CH0], c1mi[CH1];
    Nbrs<T> c0nb[CH0];
    Move<T> c0mv[CH0], c1mv[CH1];
    auto fetch0 = [&](int q, int k) {
        int r, z, y, x;
        cell0(q, r, z, y, x);
        if (y >= 0) {
            const uint32_t i = cell(y, z, x);
            c0mi[k] = load_m(mi_, NN, i);
            c0nb[k] = gather_nbrs(in, mi_, x, y, z);
            c0mv[k] = trial_move<T>(in, i, sweep, stage);
        }
    };
    auto fetch1 = [&](int q, int k) {
        int r, z, y, x;
        cell1(q, r, z, y, x);
        if (y >= 0) {
            const uint32_t i = cell(y, z, x);
            c1mi[k] = load_m(mi_, NN, i);
            c1mv[k] = trial_move<T>(in, i, sweep, stage);
        }
    };
    // spectrum of the three pairs of every layer, stored conjugated for the
    // inverse transform (as k_xinv_metro); a pair outside an open axis stays
    // zero.  Element e = ((k*3 + c)*NZ + z)*3 + p: the three pairs of one
    // (kx, component, layer) are six consecutive rows, so neighbouring lanes
    // read neighbouring 16-byte pieces of one line.
    constexpr int PER = PR * H;
    auto decode = [&](int e, int &p, int &z, int &c, int &k) {
        p = e % 3;
        const int t = e / 3;
        z = t % NZ;
        const int t2 = t / NZ;
        c = t2 % 3;
        k = t2 / 3;
    };
    // rows present of pairs p = 0, 1, 2 (bits 2p, 2p+1), the same for every
    // thread: the loads read clamped rows unconditionally and the zeros are
    // selected when the values are stored, so no load is waited for early
    unsigned okm = 0;
    #pragma unroll
    for (int p = 0; p < 3; ++p) {
        const int ya = row_of(2 * (p - 1)), yb = row_of(2 * (p - 1) + 1);
        okm |= (ya >= 0 ? 1u : 0u) << (2 * p);
        okm |= (ya >= 0 && yb >= 0 ? 1u : 0u) << (2 * p + 1);
    }
    auto spec_st = [&](int e, Pair2<cpx<T>> v) {
        int p, z, c, k;
        decode(e, p, z, c, k);
        const bool ha = (okm >> (2 * p)) & 1u, hb = (okm >> (2 * p + 1)) & 1u;
        const int rr = (p * NZ + z) * 3 + c;
        const T ax = ha ? v.a.x : T(0), ay = ha ? v.a.y : T(0);
        const T bx = hb ? v.b.x : T(0), by = hb ? v.b.y : T(0);
        cpx<T> *row = s + rr * N;
        row[k] = {ax - by, -(ay + bx)};
        if (k > 0 && N - k >= H) row[N - k] = {ax + by, ay - bx};
    };
    // the access width is fixed for the whole phase (one branch around it):
    // a per-element choice lets the compiler merge both into narrow loads
    auto load_phase = [&](auto p16c) {
        constexpr bool P16 = decltype(p16c)::value;
        auto spec_ld = [&](int e) {
            int p, z, c, k;
            decode(e, p, z, c, k);
            const int ya = row_of(2 * (p - 1)), yb = row_of(2 * (p - 1) + 1);
            const bool ha = ya >= 0, hb = ha && yb >= 0;
            const cpx<T> *col = A + c * cstride + ((long)k * NZ + z) * ny;
            Pair2<cpx<T>> v;
            if constexpr (P16) {   // complete pairs: yb = ya + 1 whenever ya >= 0
                v = load_pair16(col + (ha ? ya : 0));
            } else {
                v.a = col[ha ? ya : 0];
                v.b = col[hb ? yb : (ha ? ya : 0)];
            }
            return v;
        };
        if (PF) {
            const auto spec = staged_load<PER, NT, Pair2<cpx<T>>>(PER, spec_ld);
            #pragma unroll
            for (int q = 0; q < QC0; ++q) fetch0(q, q);
            #pragma unroll
            for (int q = 0; q < QC1; ++q) fetch1(q, q);
            staged_store<PER, NT, Pair2<cpx<T>>>(PER, spec, spec_st);
        } else {
            staged_copy<PER, NT, Pair2<cpx<T>>>(PER, spec_ld, spec_st);
        }
    };
    if constexpr (sizeof(T) == 4) {
        if (p16) load_phase(std::integral_constant<bool, true>{});
        else load_phase(std::integral_constant<bool, false>{});
    } else {
        load_phase(std::integral_constant<bool, false>{});
    }
    store_twiddles<T, N, NT>(tw, twr);
    __syncthreads();
    fft_fixed<T, N, PR, NT>(s, tw);
    // the pair's slots outside the film are zero for the forward transform
    // (no cell reads them: x >= nx, or a row past an open axis)
    const int nyt = min(2, ny - y0);
    for (int e = threadIdx.x; e < 3 * NZ * N; e += NT) {
        const int x = e % N;
        cpx<T> &v = s[3 * NZ * N + e];
        if (x >= nx) v = {T(0), T(0)};
        else if (nyt < 2) v.y = T(0);
    }
    // slot of component c at tile-relative row r (-2..3), layer z, column x;
    // an odd row sits in the imaginary part (its demag value negated)
    auto slot = [&](int r, int z, int c, int x) -> cpx<T> & {
        return s[((((r + 2) >> 1) * NZ + z) * 3 + c) * N + x];
    };
    auto hval = [&](int r, int z, int x) -> V3<T> {
        if ((r + 2) & 1) return {-slot(r, z, 0, x).y, -slot(r, z, 1, x).y, -slot(r, z, 2, x).y};
        return {slot(r, z, 0, x).x, slot(r, z, 1, x).x, slot(r, z, 2, x).x};
    };
    auto put = [&](int r, int z, int x, const V3<T> &v) {
        if ((r + 2) & 1) {
            slot(r, z, 0, x).y = v.x;
            slot(r, z, 1, x).y = v.y;
            slot(r, z, 2, x).y = v.z;
        } else {
            slot(r, z, 0, x).x = v.x;
            slot(r, z, 1, x).x = v.y;
            slot(r, z, 2, x).x = v.z;
        }
    };
    auto mval = [&](int r, int z, int x) -> V3<T> {   // new value written by put()
        if ((r + 2) & 1) return {slot(r, z, 0, x).y, slot(r, z, 1, x).y, slot(r, z, 2, x).y};
        return {slot(r, z, 0, x).x, slot(r, z, 1, x).x, slot(r, z, 2, x).x};
    };
    // colour 0 on rows -1..2 (neighbours: pre-sweep colour-1 cells of m_in)
    #pragma unroll
    for (int q0 = 0; q0 < QC0; q0 += CH0) {
        if (!PF) {
            #pragma unroll
            for (int k = 0; k < CH0; ++k) fetch0(q0 + k, k);
        }
        #pragma unroll
        for (int k = 0; k < CH0; ++k) {
            int r, z, y, x;
            cell0(q0 + k, r, z, y, x);
            if (y < 0) continue;
            V3<T> ex, dmi;
            bond_sums(in, c0nb[k].xp, c0nb[k].xm, c0nb[k].yp, c0nb[k].ym, c0nb[k].zp, c0nb[k].zm,
                      ex, dmi);
            const V3<T> mn = metro_accept(in, c0mi[k], ex, dmi, hval(r, z, x), 1, c0mv[k], bt)
                                 ? c0mv[k].mn
                                 : c0mi[k];
            put(r, z, x, mn);
            if (r == 0 || r == 1) {
                const long i = cell(y, z, x);
                mo[i] = mn.x;
                mo[NN + i] = mn.y;
                mo[2 * NN + i] = mn.z;
            }
        }
    }
    __syncthreads();
    // colour 1 on rows 0, 1 (neighbours: the colour-0 values just computed)
    {
        const V3<T> zero = {T(0), T(0), T(0)};
        #pragma unroll
        for (int q0 = 0; q0 < QC1; q0 += CH1) {
            if (!PF) {
                #pragma unroll
                for (int k = 0; k < CH1; ++k) fetch1(q0 + k, k);
            }
            #pragma unroll
            for (int k = 0; k < CH1; ++k) {
                int r, z, y, x;
                cell1(q0 + k, r, z, y, x);
                if (y < 0) continue;
                Nbrs<T> nb = {zero, zero, zero, zero, zero, zero};
                int t;
                if ((t = nbr(x, +1, nx, in.pbc[0])) >= 0) nb.xp = mval(r, z, t);
                if ((t = nbr(x, -1, nx, in.pbc[0])) >= 0) nb.xm = mval(r, z, t);
                if (nbr(y, +1, ny, pery) >= 0) nb.yp = mval(r + 1, z, x);
                if (nbr(y, -1, ny, pery) >= 0) nb.ym = mval(r - 1, z, x);
                if (NZ > 1) {
                    if ((t = nbr(z, +1, NZ, in.pbc[2])) >= 0) nb.zp = mval(r, t, x);
                    if ((t = nbr(z, -1, NZ, in.pbc[2])) >= 0) nb.zm = mval(r, t, x);
                }
                V3<T> ex, dmi;
                bond_sums(in, nb.xp, nb.xm, nb.yp, nb.ym, nb.zp, nb.zm, ex, dmi);
                const V3<T> mn = metro_accept(in, c1mi[k], ex, dmi, hval(r, z, x), 1, c1mv[k], bt)
                                     ? c1mv[k].mn
                                     : c1mi[k];
                const long i = cell(y, z, x);
                mo[i] = mn.x;
                mo[NN + i] = mn.y;
                mo[2 * NN + i] = mn.z;
                put(r, z, x, mn);
            }
        }
    }
    __syncthreads();
    // the pair's final state (zero outside the film) forward into A
    fft_fixed<T, N, 3 * NZ, NT>(s + 3 * NZ * N, tw);
    for (int e = threadIdx.x; e < 3 * NZ * H; e += NT) {
        const int rr = e / H, k = e - rr * H;
        const int z = rr / 3, c = rr - 3 * z;
        const cpx<T> *row = s + (3 * NZ + rr) * N;
        const cpx<T> Z = row[k];
        const cpx<T> B = row[k ? N - k : 0];
        cpx<T> *o = A + c * cstride + ((long)k * NZ + z) * ny + y0;
        const cpx<T> oa = {(Z.x + B.x) * T(0.5), (Z.y - B.y) * T(0.5)};
        const cpx<T> ob = {(Z.y + B.y) * T(0.5), (B.x - Z.x) * T(0.5)};
        if (p16) {
            store_pair16(o, oa, ob);
        } else {
            o[0] = oa;
            if (nyt > 1) o[1] = ob;
        }
    }
}

template <typename T>
hipError_t launch_row_sweep(hipStream_t st, const Inst *I, const Work *W, int n, int fixed_n, int nz,
                            cpx<T> *spec, const T *m_in, T *m_out, const double *beta,
                            const uint32_t *ctl, uint32_t sweep_off, const cpx<T> *tw, size_t lds)
{
    if (n <= 0) return hipSuccess;
    const dim3 g(n);
#define MMC_RS_CASE(NN, ZZ)                                                                      \
    if (fixed_n == NN && nz == ZZ) {                                                             \
        MMC_LAUNCH((k_row_sweep<T, NN, ZZ>), g, dim3(kRowThreads), lds, st, I, W, spec, m_in,    \
                   m_out, beta, ctl, sweep_off, tw);                                             \
        return post();                                                                           \
    }
#define MMC_RS_N(NN) MMC_RS_CASE(NN, 1) MMC_RS_CASE(NN, 2)
    MMC_RS_N(128) MMC_RS_N(192) MMC_RS_N(256) MMC_RS_N(384) MMC_RS_N(512) MMC_RS_N(1024)
    MMC_RS_N(2048)
#undef MMC_RS_N
#undef MMC_RS_CASE
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t set_lds_rows(int bytes)
{
    hipError_t e = hipSuccess;
#define MMC_RSL(NN) set_lds_one(e, (const void *)k_row_sweep<T, NN, 1>, bytes); \
    set_lds_one(e, (const void *)k_row_sweep<T, NN, 2>, bytes);
    MMC_RSL(128) MMC_RSL(192) MMC_RSL(256) MMC_RSL(384) MMC_RSL(512) MMC_RSL(1024) MMC_RSL(2048)
#undef MMC_RSL
    return e;
}

#define MMC_INST(T)                                                                           \
    template hipError_t launch_row_sweep<T>(hipStream_t, const Inst *, const Work *, int, int,  \
                                            int, cpx<T> *, const T *, T *, const double *,   \
                                            const uint32_t *, uint32_t, const cpx<T> *,      \
                                            size_t);                                         \
    template hipError_t set_lds_rows<T>(int);
MMC_INST(float)
MMC_INST(double)
#undef MMC_INST

}  // namespace mmc
